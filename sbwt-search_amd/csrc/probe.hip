// Random-line probe: the ceiling of dependent random 128-byte line reads on
// this GPU at a given working-set size, against which the search kernels'
// line rates are measured (their cost unit is one 128-byte line per random
// access). No counterpart in the reference.
#include <algorithm>

#include "probe.hpp"

namespace sbwt_amd {

namespace {

// A full-period cycle over 2^bits lines: line i holds next(i) = (a i + c) mod
// 2^bits (Hull-Dobell: c odd, a = 1 mod 4), so each load's value is the next
// load's address and every chain is a true dependence chain.
constexpr u64 kProbeA = 6364136223846793005ULL, kProbeC = 1442695040888963407ULL;

__global__ __launch_bounds__(256) void k_probe_fill(uint4 *table, u64 lines, u64 mask) {
  for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < lines;
       i += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 next = (kProbeA * i + kProbeC) & mask;
    uint4 *line = table + i * 8;
    line[0] = make_uint4(static_cast<u32>(next), static_cast<u32>(next >> 32), 0, 0);
  }
}

// Each lane follows its own chain for `steps` lines from a scattered start;
// one 16-byte load per line. The sum keeps the chains live.
__global__ __launch_bounds__(256) void k_probe_chase(const uint4 *__restrict__ table, u64 mask,
                                                     u32 steps, u64 *__restrict__ sink) {
  const u64 lane = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
  u64 i = (lane * 0x9E3779B97F4A7C15ULL) & mask;
  u64 acc = 0;
  for (u32 s = 0; s < steps; ++s) {
    const uint4 v = table[i * 8];
    i = (static_cast<u64>(v.x) | (static_cast<u64>(v.y) << 32)) & mask;
    acc += i;
  }
  if (acc == 0x5A5A5A5A5A5A5A5AULL) sink[0] = acc;  // never true in practice; keeps the loads
}

}  // namespace

ProbeResult random_line_probe(u64 working_set_bytes, u32 blocks_per_cu, u32 steps) {
  if (working_set_bytes < (1u << 20)) throw std::runtime_error("probe working set below 1 MiB");
  u32 bits = 0;
  while ((128ULL << bits) < working_set_bytes) ++bits;  // lines: smallest 2^bits that holds it
  const u64 lines = 1ULL << bits, mask = lines - 1;
  int dev = 0, cus = 1;
  SBWT_HIP(hipGetDevice(&dev));
  SBWT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  uint4 *table = nullptr;
  u64 *sink = nullptr;
  hipEvent_t ev[2];
  SBWT_HIP(hipMalloc(&table, lines * 128));
  SBWT_HIP(hipMalloc(&sink, 8));
  for (auto &e : ev) SBWT_HIP(hipEventCreate(&e));
  ProbeResult r{};
  try {
    k_probe_fill<<<static_cast<u32>(std::min<u64>((lines + 255) / 256, 1u << 20)), 256>>>(table, lines, mask);
    SBWT_HIP(hipGetLastError());
    const u32 grid = static_cast<u32>(std::max(1, cus)) * std::max(1u, blocks_per_cu);
    // one untimed pass warms the caches to their steady state
    k_probe_chase<<<grid, 256>>>(table, mask, steps, sink);
    SBWT_HIP(hipEventRecord(ev[0], nullptr));
    k_probe_chase<<<grid, 256>>>(table, mask, steps, sink);
    SBWT_HIP(hipEventRecord(ev[1], nullptr));
    SBWT_HIP(hipGetLastError());
    SBWT_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    SBWT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    r.ms = ms;
    r.lines = static_cast<u64>(grid) * 256 * steps;
    r.table_bytes = lines * 128;
  } catch (...) {
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipFree(table);
    (void)hipFree(sink);
    throw;
  }
  for (auto &e : ev) SBWT_HIP(hipEventDestroy(e));
  SBWT_HIP(hipFree(table));
  SBWT_HIP(hipFree(sink));
  return r;
}

}  // namespace sbwt_amd
