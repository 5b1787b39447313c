// Device-wide inclusive scan of u64 values (see scan.hpp).
#include <algorithm>

#include "scan.hpp"

namespace sbwt_amd {

namespace {

constexpr u32 kScanThreads = 256;
constexpr u32 kScanItems = 16;
constexpr u64 kScanChunk = static_cast<u64>(kScanThreads) * kScanItems;

template <u32 N>
__device__ __forceinline__ u64 block_inclusive_scan(u64 v, u64 *sh) {
  const u32 t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (u32 off = 1; off < N; off <<= 1) {
    const u64 x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const u64 r = sh[t];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const u64 *__restrict__ in, u64 n,
                                                               u64 *__restrict__ sums) {
  __shared__ u64 sh[kScanThreads];
  const u64 first = blockIdx.x * kScanChunk + threadIdx.x * static_cast<u64>(kScanItems);
  u64 s = 0;
  for (u32 i = 0; i < kScanItems; ++i)
    if (first + i < n) s += in[first + i];
  const u64 tot = block_inclusive_scan<kScanThreads>(s, sh);
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_chunks(u64 *__restrict__ sums, u64 n_chunks) {
  __shared__ u64 sh[1024];
  __shared__ u64 tile_total;
  u64 carry = 0;
  for (u64 base = 0; base < n_chunks; base += 1024) {
    const u64 i = base + threadIdx.x;
    const u64 v = i < n_chunks ? sums[i] : 0;
    const u64 inc = block_inclusive_scan<1024>(v, sh);
    if (i < n_chunks) sums[i] = carry + inc - v;
    if (threadIdx.x == 1023) tile_total = inc;
    __syncthreads();
    carry += tile_total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_write(const u64 *__restrict__ in, u64 n,
                                                              const u64 *__restrict__ offs,
                                                              u64 *__restrict__ out) {
  __shared__ u64 sh[kScanThreads];
  const u64 first = blockIdx.x * kScanChunk + threadIdx.x * static_cast<u64>(kScanItems);
  u64 v[kScanItems];
  u64 s = 0;
  for (u32 i = 0; i < kScanItems; ++i) {
    v[i] = first + i < n ? in[first + i] : 0;
    s += v[i];
  }
  const u64 inc = block_inclusive_scan<kScanThreads>(s, sh);
  u64 run = offs[blockIdx.x] + inc - s;
  for (u32 i = 0; i < kScanItems; ++i) {
    run += v[i];
    if (first + i < n) out[first + i] = run;
  }
}

}  // namespace

u64 scan_tmp_entries(u64 n) { return n / kScanChunk + 2; }

void scan_inclusive_u64(const u64 *d_in, u64 n, u64 *d_out, u64 *d_tmp, hipStream_t s) {
  if (n == 0) return;
  const u64 chunks = (n + kScanChunk - 1) / kScanChunk;
  k_scan_reduce<<<static_cast<u32>(chunks), kScanThreads, 0, s>>>(d_in, n, d_tmp);
  k_scan_chunks<<<1, 1024, 0, s>>>(d_tmp, chunks);
  k_scan_write<<<static_cast<u32>(chunks), kScanThreads, 0, s>>>(d_in, n, d_tmp, d_out);
  SBWT_HIP(hipGetLastError());
}

}  // namespace sbwt_amd
