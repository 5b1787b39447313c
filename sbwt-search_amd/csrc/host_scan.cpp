// Counting loops written so that the compiler vectorizes them; built for
// AVX2 and for the baseline x86-64, picked at load time.
#include "host_scan.hpp"

#include <cstring>

namespace sbwt_amd {

__attribute__((target_clones("avx2", "default"))) u64 count_byte(const char *p, u64 n, char ch) {
  const auto *u = reinterpret_cast<const unsigned char *>(p);
  const auto c = static_cast<unsigned char>(ch);
  constexpr u64 kLanes = 32, kRounds = 255;  // byte counters cannot overflow in 255 rounds
  u64 total = 0, i = 0;
  while (i + kLanes * kRounds <= n) {
    unsigned char acc[kLanes] = {};
    for (u64 r = 0; r < kRounds; ++r, i += kLanes)
      for (u64 k = 0; k < kLanes; ++k) acc[k] = static_cast<unsigned char>(acc[k] + (u[i + k] == c));
    for (u64 k = 0; k < kLanes; ++k) total += acc[k];
  }
  for (; i < n; ++i) total += u[i] == c;
  return total;
}

__attribute__((target_clones("avx2", "default"))) u64 count_word(const char *p, u64 n_words, u64 v) {
  u64 total = 0;
  for (u64 i = 0; i < n_words; ++i) {
    u64 w;
    std::memcpy(&w, p + 8 * i, 8);
    total += w == v;
  }
  return total;
}

}  // namespace sbwt_amd
