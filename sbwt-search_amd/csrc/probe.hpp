// Random-line probe (see probe.hip).
#pragma once

#include "common.hpp"

namespace sbwt_amd {

struct ProbeResult {
  float ms;         // one timed pass
  u64 lines;        // 128-byte lines read in it
  u64 table_bytes;  // working set actually used (2^b lines of 128 bytes)
};

// Dependent random 16-byte loads, one per 128-byte line, over a table of the
// smallest power-of-two number of lines that holds working_set_bytes, from
// blocks_per_cu 256-lane workgroups per CU each following `steps` lines.
ProbeResult random_line_probe(u64 working_set_bytes, u32 blocks_per_cu, u32 steps);

}  // namespace sbwt_amd
