// Parallel scan of a byte automaton with at most four states, shared by the
// device parsers of text inputs. Each thread takes a 16-byte segment, each
// workgroup of 256 threads a 4 KiB block:
//   1. a block's state map (next state for each start state, 2 bits each,
//      composed over its bytes in order) is reduced per block;
//   2. one workgroup scans the block maps into each block's start state;
//   3. a block scans its segment maps, so every thread knows the state its
//      segment starts in, and walks its 16 bytes from there.
// The automaton is given as a per-byte map function; what a walk emits is
// the parser's.
#pragma once

#include "common.hpp"

namespace sbwt_amd {
namespace automaton {

constexpr u32 kThreads = 256;
constexpr u32 kSegBytes = 16;
constexpr u64 kBlockBytes = u64{kThreads} * kSegBytes;
constexpr u32 kStatesThreads = 1024;
constexpr u32 kMapIdentity = 0 | 1 << 2 | 2 << 4 | 3 << 6;

__device__ __forceinline__ u32 next_state(u32 map, u32 s) { return (map >> (2 * s)) & 3u; }

// f, then g.
__device__ __forceinline__ u32 compose(u32 f, u32 g) {
  u32 h = 0;
#pragma unroll
  for (u32 s = 0; s < 4; ++s) h |= next_state(g, next_state(f, s)) << (2 * s);
  return h;
}

// The bytes of segment `seg` (zero past n); returns how many are in range.
// `body` is 16-byte aligned and readable in whole 16-byte chunks.
__device__ __forceinline__ u32 load_segment(const u8 *body, u64 n, u64 seg, u8 (&b)[kSegBytes]) {
  const u64 beg = seg * kSegBytes;
  if (beg >= n) return 0;
  const u32 len = static_cast<u32>(min<u64>(kSegBytes, n - beg));
  const uint4 v = *reinterpret_cast<const uint4 *>(body + beg);
  const u32 w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (u32 i = 0; i < kSegBytes; ++i) b[i] = i < len ? static_cast<u8>(w[i / 4] >> (8 * (i % 4))) : 0;
  return len;
}

template <typename CharMap>
__device__ __forceinline__ u32 segment_map(const u8 (&b)[kSegBytes], u32 len, CharMap char_map) {
  u32 m = kMapIdentity;
#pragma unroll
  for (u32 i = 0; i < kSegBytes; ++i)
    if (i < len) m = compose(m, char_map(b[i]));
  return m;
}

// Exclusive prefix (in thread order) of the maps of a workgroup of WAVES
// waves; *total gets the composition of all of them.
template <u32 WAVES>
__device__ __forceinline__ u32 block_scan_maps(u32 m, u32 *sh, u32 *total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32 incl = m;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(incl, d, 64);
    if (lane >= d) incl = compose(y, incl);
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  u32 before = kMapIdentity, all = kMapIdentity;
#pragma unroll
  for (u32 w = 0; w < WAVES; ++w) {
    if (w < wave) before = compose(before, sh[w]);
    all = compose(all, sh[w]);
  }
  __syncthreads();
  u32 excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = kMapIdentity;
  *total = all;
  return compose(before, excl);
}

// Workgroup sum (every thread gets it) and exclusive prefix of x.
__device__ __forceinline__ u32 block_prefix_sum(u32 x, u32 *sh, u32 *total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32 incl = x;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  u32 before = 0, all = 0;
#pragma unroll
  for (u32 w = 0; w < kThreads / 64; ++w) {
    if (w < wave) before += sh[w];
    all += sh[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

// Step 1: maps[blk] = composed map of block blk.
template <typename CharMap>
__device__ __forceinline__ void block_maps(const u8 *body, u64 n, u64 n_blocks, u8 *maps,
                                           CharMap char_map) {
  __shared__ u32 sh[kThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(body, n, blk * kThreads + threadIdx.x, b);
    u32 total;
    block_scan_maps<kThreads / 64>(segment_map(b, len, char_map), sh, &total);
    if (threadIdx.x == 0) maps[blk] = static_cast<u8>(total);
  }
}

// Step 2 (one workgroup of kStatesThreads): states[blk] = state at the start
// of block blk, the body starting in state `first`.
__device__ __forceinline__ void block_states(const u8 *maps, u64 n_blocks, u32 first, u8 *states) {
  __shared__ u32 sh[kStatesThreads / 64];
  const u64 per = (n_blocks + kStatesThreads - 1) / kStatesThreads;
  const u64 beg = min<u64>(n_blocks, threadIdx.x * per), end = min<u64>(n_blocks, beg + per);
  u32 m = kMapIdentity;
  for (u64 i = beg; i < end; ++i) m = compose(m, maps[i]);
  u32 total;
  const u32 before = block_scan_maps<kStatesThreads / 64>(m, sh, &total);
  u32 st = next_state(before, first);
  for (u64 i = beg; i < end; ++i) {
    states[i] = static_cast<u8>(st);
    st = next_state(maps[i], st);
  }
}

// Step 3: state at the start of this thread's segment (bytes b[0, len)).
template <typename CharMap>
__device__ __forceinline__ u32 segment_start(u64 blk, const u8 *states, u32 *sh,
                                             const u8 (&b)[kSegBytes], u32 len, CharMap char_map) {
  u32 total;
  const u32 before = block_scan_maps<kThreads / 64>(segment_map(b, len, char_map), sh, &total);
  return next_state(before, states[blk]);
}

}  // namespace automaton
}  // namespace sbwt_amd
