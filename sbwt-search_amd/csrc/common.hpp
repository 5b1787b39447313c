// Shared definitions of the MI355X (gfx950) search library: error handling,
// device-side index views and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace sbwt_amd {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr u64 kNotFound = ~0ULL;        // reference "-1" (IndexSearcher.cuh:57)
constexpr u64 kInvalid = ~0ULL - 1;     // reference "-2" (BinaryContinuousIndexResultsPrinter)
constexpr u64 kWalkStuck = ~0ULL - 3;   // key walk reached a row with no edge and no mark

// HIP failures become std::runtime_error("GPUassert: <msg> at <file>:<line>"),
// the message shape of the reference's GPU_CHECK (src/Tools/GpuUtils.h:20-29).
[[noreturn]] void throw_hip(hipError_t e, const char *file, int line);
#define SBWT_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) ::sbwt_amd::throw_hip(e_, __FILE__, __LINE__); \
  } while (0)

// ---------------------------------------------------------------------------
// Device layout of the SBWT ("ACGT rank blocks").
// Block b covers nodes [96 b, 96 b + 96) and is one 64-byte line holding, for
// each letter c, a 16-byte entry {count, bits[0..31], bits[32..63],
// bits[64..95]}: count = number of 1s of letter c before the block, relative
// to the block's hyperblock (2^25 blocks), so rank_c(i) is one 16-byte load
// plus one small table read. The key-k-mer marks (and the color index's
// marks) use 64-byte "mark blocks": {u64 count, 7 words} for 448 bits.
// ---------------------------------------------------------------------------
constexpr u32 kBlockNodes = 96;
constexpr u32 kHyperShift = 25;  // blocks per hyperblock = 2^25 (3.2 G nodes)
constexpr u32 kMarkBits = 448;

struct IndexView {
  const uint4 *blocks;        // [n_blocks][4] entries
  const u64 *hyper_base;      // [n_hyper][4] absolute count at hyperblock start
  const void *presearch;      // [4^kp] of {left,right} (uint2 or ulonglong2)
  const u64 *key_marks;       // mark blocks of the key-k-mer marks, or null
  const u64 *sgs;             // suffix-group starts (plain bit vector), or null
  u64 c_map[5];
  u64 num_nodes;
  u32 k;
  u32 kp;
};

// Threads per workgroup of the search kernels: 4 wave64s.
constexpr u32 kSearchThreads = 256;

}  // namespace sbwt_amd
