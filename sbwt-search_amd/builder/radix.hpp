// Parallel LSD radix sort of u64 keys (optionally carrying a u32 payload),
// 11 bits per pass over the low `key_bits` bits. OpenMP, stable.
#pragma once
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace sbwt_build {

template <bool with_payload>
inline void radix_sort_impl(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, int threads) {
  if (n < 2) return;
  constexpr int kBits = 11;
  constexpr uint64_t kBuckets = 1ULL << kBits;
  std::vector<uint64_t> ktmp(n);
  std::vector<uint32_t> vtmp(with_payload ? n : 0);
  uint64_t *src_k = keys, *dst_k = ktmp.data();
  uint32_t *src_v = vals, *dst_v = with_payload ? vtmp.data() : nullptr;
  const int T = threads > 0 ? threads : 1;
  std::vector<uint64_t> counts(static_cast<size_t>(T) * kBuckets);
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += kBits, ++passes) {
    std::fill(counts.begin(), counts.end(), 0);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
      uint64_t *c = counts.data() + static_cast<size_t>(t) * kBuckets;
      for (uint64_t i = lo; i < hi; ++i) c[(src_k[i] >> shift) & (kBuckets - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        uint64_t sum = 0;
        for (uint64_t b = 0; b < kBuckets; ++b)
          for (int tt = 0; tt < T; ++tt) {
            uint64_t &x = counts[static_cast<size_t>(tt) * kBuckets + b];
            const uint64_t v = x;
            x = sum;
            sum += v;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) {
        const uint64_t pos = c[(src_k[i] >> shift) & (kBuckets - 1)]++;
        dst_k[pos] = src_k[i];
        if (with_payload) dst_v[pos] = src_v[i];
      }
    }
    std::swap(src_k, dst_k);
    if (with_payload) std::swap(src_v, dst_v);
  }
  if (passes % 2) {
    std::memcpy(keys, src_k, n * 8);
    if (with_payload) std::memcpy(vals, src_v, n * 4);
  }
}

inline void radix_sort(uint64_t *keys, uint64_t n, int key_bits, int threads) {
  radix_sort_impl<false>(keys, nullptr, n, key_bits, threads);
}
inline void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, int threads) {
  radix_sort_impl<true>(keys, vals, n, key_bits, threads);
}

}  // namespace sbwt_build
