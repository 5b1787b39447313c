// Device index construction: ACGT rank blocks and mark blocks (built on the
// host, uploaded once) and the presearch table (built on the device).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "device_fns.hpp"
#include "device_index.hpp"

namespace sbwt_amd {

void throw_hip(hipError_t e, const char *file, int line) {
  throw std::runtime_error(std::string("GPUassert: ") + hipGetErrorString(e) + " at " + file + ":" +
                           std::to_string(line));
}

u32 default_presearch_letters(u32 k) { return k < 12 ? k : 12; }

namespace {

template <typename F>
void parallel_for(u64 n, F f) {
  unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) T = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t)
    th.emplace_back([=, &f] {
      for (u64 i = n * t / T; i < n * (t + 1) / T; ++i) f(i);
    });
  for (auto &x : th) x.join();
}

// `len` (<= 64) bits of a bit vector starting at bit `start`, zero past `n_bits`.
inline u64 get_bits(const u64 *w, u64 n_bits, u64 start, u32 len) {
  if (start >= n_bits || len == 0) return 0;
  const u64 wi = start >> 6;
  const u32 off = static_cast<u32>(start & 63);
  const u64 n_words = (n_bits + 63) >> 6;
  u64 x = w[wi] >> off;
  if (off && wi + 1 < n_words) x |= w[wi + 1] << (64 - off);
  u64 avail = n_bits - start;
  u32 keep = static_cast<u32>(std::min<u64>(len, avail));
  return keep >= 64 ? x : (x & ((1ULL << keep) - 1));
}

template <typename T>
T *upload(const T *host, u64 count) {
  T *d = nullptr;
  SBWT_HIP(hipMalloc(&d, std::max<u64>(count, 1) * sizeof(T)));
  if (count) SBWT_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// Presearch: interval of every kp-letter prefix x (first letter in the top two
// of its 2*kp bits), src/Presearcher/Presearcher.cuh:19-41.
template <typename Idx, typename Pair>
__global__ __launch_bounds__(256) void k_presearch(IndexView ix, u32 kp, Pair *out, u64 count) {
  for (u64 x = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; x < count;
       x += static_cast<u64>(gridDim.x) * blockDim.x) {
    u32 c = static_cast<u32>(x >> (2 * kp - 2)) & 3u;
    Idx l = static_cast<Idx>(ix.c_map[c]);
    Idx r = static_cast<Idx>(ix.c_map[c + 1] - 1);
    for (int sh = 2 * static_cast<int>(kp) - 4; sh >= 0; sh -= 2) {
      c = static_cast<u32>(x >> sh) & 3u;
      sbwt_step<Idx>(ix, c, l, r);
    }
    Pair p;
    p.x = l;
    p.y = r;
    out[x] = p;
  }
}

// Longer prefixes from a base table of kb-letter intervals: the interval of
// x is that of its first kb letters stepped by the remaining ones. Stepping
// stops once the interval is empty (l > r), which every later step keeps.
template <typename Idx, typename Pair>
__global__ __launch_bounds__(256) void k_presearch_extend(IndexView ix, const Pair *__restrict__ base,
                                                          u32 kb, u32 kp, Pair *out, u64 count) {
  const u32 extra = 2 * (kp - kb);
  for (u64 x = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; x < count;
       x += static_cast<u64>(gridDim.x) * blockDim.x) {
    const Pair b = base[x >> extra];
    Idx l = static_cast<Idx>(b.x), r = static_cast<Idx>(b.y);
    for (int sh = static_cast<int>(extra) - 2; sh >= 0 && l <= r; sh -= 2)
      sbwt_step<Idx>(ix, static_cast<u32>(x >> sh) & 3u, l, r);
    Pair p;
    p.x = l;
    p.y = r;
    out[x] = p;
  }
}

template <typename Idx, typename Pair>
void build_presearch(const IndexView &ix, u32 kp, Pair *out) {
  constexpr u32 kBase = 12;
  const u64 count = 1ULL << (2 * kp);
  const u32 grid = static_cast<u32>(std::min<u64>((count + 255) / 256, 1u << 20));
  if (kp <= kBase) {
    k_presearch<Idx, Pair><<<grid, 256>>>(ix, kp, out, count);
    SBWT_HIP(hipGetLastError());
    return;
  }
  const u64 base_count = 1ULL << (2 * kBase);
  Pair *base = nullptr;
  SBWT_HIP(hipMalloc(&base, base_count * sizeof(Pair)));
  k_presearch<Idx, Pair><<<static_cast<u32>(base_count / 256), 256>>>(ix, kBase, base, base_count);
  SBWT_HIP(hipGetLastError());
  k_presearch_extend<Idx, Pair><<<grid, 256>>>(ix, base, kBase, kp, out, count);
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipDeviceSynchronize());
  SBWT_HIP(hipFree(base));
}

}  // namespace

void DeviceMarks::build(const u64 *words, u64 bits) {
  n_bits = bits;
  n_blocks = bits / kMarkBits + 1;  // rank(n_bits) stays in range
  std::vector<u64> host(n_blocks * 8, 0);
  parallel_for(n_blocks, [&](u64 b) {
    for (u32 w = 0; w < 7; ++w) host[b * 8 + 1 + w] = get_bits(words, bits, b * kMarkBits + 64 * w, 64);
  });
  u64 run = 0;
  for (u64 b = 0; b < n_blocks; ++b) {
    host[b * 8] = run;
    for (u32 w = 0; w < 7; ++w) run += __builtin_popcountll(host[b * 8 + 1 + w]);
  }
  total = run;
  blocks = upload(host.data(), host.size());
}

void DeviceMarks::release() {
  if (blocks) (void)hipFree(blocks);
  blocks = nullptr;
}

void DeviceIndex::build(const u64 *const acgt[4], u64 n_, u32 k_, u32 kp_, const u64 *km_words,
                        u64 km_bits, const u64 *sgs_words) {
  SBWT_HIP(hipGetDevice(&device));
  n = n_;
  k = k_;
  kp = kp_;
  // 64-bit node indexes are needed from 2^32 nodes on; SBWT_AMD_FORCE_WIDE=1
  // selects them for any index so the 64-bit kernels can be validated small.
  const char *force = std::getenv("SBWT_AMD_FORCE_WIDE");
  wide = n >= (1ULL << 32) || (force && force[0] == '1');
  n_blocks = n / kBlockNodes + 1;  // the block of node n exists: rank(n) is defined
  // per-block popcounts, then per-letter prefix sums
  std::vector<uint4> host(n_blocks * 4);
  parallel_for(n_blocks, [&](u64 b) {
    for (u32 c = 0; c < 4; ++c) {
      const u64 lo = get_bits(acgt[c], n, b * kBlockNodes, 64);
      const u64 hi = get_bits(acgt[c], n, b * kBlockNodes + 64, 32);
      uint4 e;
      e.x = 0;
      e.y = static_cast<u32>(lo);
      e.z = static_cast<u32>(lo >> 32);
      e.w = static_cast<u32>(hi);
      host[b * 4 + c] = e;
    }
  });
  n_hyper = (n_blocks >> kHyperShift) + 1;
  std::vector<u64> hyper(n_hyper * 4, 0);
  u64 run[4] = {0, 0, 0, 0};
  for (u64 b = 0; b < n_blocks; ++b) {
    for (u32 c = 0; c < 4; ++c) {
      if ((b & ((1ULL << kHyperShift) - 1)) == 0) hyper[(b >> kHyperShift) * 4 + c] = run[c];
      uint4 &e = host[b * 4 + c];
      e.x = static_cast<u32>(wide ? run[c] - hyper[(b >> kHyperShift) * 4 + c] : run[c]);
      run[c] += __builtin_popcount(e.y) + __builtin_popcount(e.z) + __builtin_popcount(e.w);
    }
  }
  c_map[0] = 1;  // src/SbwtBuilder/SbwtBuilder.cpp:111,128-130
  for (u32 c = 0; c < 4; ++c) c_map[c + 1] = c_map[c] + run[c];
  blocks = upload(host.data(), host.size());
  hyper_base = upload(hyper.data(), hyper.size());
  device_bytes = host.size() * sizeof(uint4) + hyper.size() * 8;
  host.clear();
  host.shrink_to_fit();
  if (km_words) {
    key_marks.build(km_words, km_bits);
    device_bytes += key_marks.n_blocks * 64;
  }
  if (sgs_words) {
    std::vector<u64> w((n + 63) / 64 + 2, 0);
    for (u64 i = 0; i < (n + 63) / 64; ++i) w[i] = sgs_words[i];
    if (n % 64) w[n / 64] &= (1ULL << (n % 64)) - 1;
    w[0] |= 1;  // the root starts the first group
    sgs = upload(w.data(), w.size());
    device_bytes += w.size() * 8;
  }
  // presearch table on the device
  const u64 count = 1ULL << (2 * kp);
  const IndexView ix = view();
  if (wide) {
    presearch_bytes = count * sizeof(ulonglong2);
    SBWT_HIP(hipMalloc(&presearch, presearch_bytes));
    build_presearch<u64, ulonglong2>(ix, kp, static_cast<ulonglong2 *>(presearch));
  } else {
    presearch_bytes = count * sizeof(uint2);
    SBWT_HIP(hipMalloc(&presearch, presearch_bytes));
    build_presearch<u32, uint2>(ix, kp, static_cast<uint2 *>(presearch));
  }
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipDeviceSynchronize());
  device_bytes += presearch_bytes;
}

void DeviceIndex::release() {
  if (blocks) (void)hipFree(blocks);
  if (hyper_base) (void)hipFree(hyper_base);
  if (presearch) (void)hipFree(presearch);
  if (sgs) (void)hipFree(sgs);
  sgs = nullptr;
  blocks = nullptr;
  hyper_base = nullptr;
  presearch = nullptr;
  key_marks.release();
}

IndexView DeviceIndex::view() const {
  IndexView v{};
  v.blocks = blocks;
  v.hyper_base = hyper_base;
  v.presearch = presearch;
  v.key_marks = key_marks.blocks;
  v.sgs = sgs;
  for (int i = 0; i < 5; ++i) v.c_map[i] = c_map[i];
  v.num_nodes = n;
  v.k = k;
  v.kp = kp;
  return v;
}

}  // namespace sbwt_amd
