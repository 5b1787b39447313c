// Device-wide inclusive scan of u64 values (three kernels: per-chunk sums,
// one-workgroup scan of the sums, per-chunk scan + offset).
#pragma once

#include "common.hpp"

namespace sbwt_amd {

// Entries of the temporary buffer a scan of n values needs.
u64 scan_tmp_entries(u64 n);
void scan_inclusive_u64(const u64 *d_in, u64 n, u64 *d_out, u64 *d_tmp, hipStream_t s);

}  // namespace sbwt_amd
