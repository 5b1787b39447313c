// Synthetic workloads, SBWT builder/writer and color-index builder/writer.
// See builder.h for the contract of every function.
#include "builder.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "builder_impl.hpp"
#include "radix.hpp"

using sbwt_build::radix_sort;
using sbwt_build::radix_sort_pairs;
using sbwt_build::bits_for;
using sbwt_build::low_mask;
using std::vector;

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double unit(uint64_t x) { return static_cast<double>(x >> 11) * 0x1.0p-53; }

struct Rng {
  uint64_t base, i = 0;
  Rng(uint64_t seed, uint64_t stream) : base(mix64(seed * 0xD1B54A32D192ED03ULL + stream)) {}
  uint64_t next() { return mix64(base + i++); }
  double uniform() { return unit(next()); }
};

inline int threads_or_default(int t) { return t > 0 ? t : omp_get_max_threads(); }

// ---- binary output helpers (sdsl-lite c32874c on-disk forms) ----
struct Out {
  FILE *f;
  bool ok = true;
  void raw(const void *p, size_t n) {
    if (n && fwrite(p, 1, n, f) != n) ok = false;
  }
  void u64(uint64_t v) { raw(&v, 8); }
  void str(const std::string &s) {
    u64(s.size());
    raw(s.data(), s.size());
  }
  // sdsl bit_vector: u64 size in bits, then ceil(bits/64) words
  void bit_vector(const vector<uint64_t> &w, uint64_t bits) {
    u64(bits);
    raw(w.data(), (bits + 63) / 64 * 8);
  }
  // sdsl int_vector<> (runtime width): u64 size in bits, u8 width, words
  void int_vector(const vector<uint64_t> &vals, int width) {
    const uint64_t bits = vals.size() * static_cast<uint64_t>(width);
    vector<uint64_t> w((bits + 63) / 64 + 1, 0);
    for (uint64_t i = 0; i < vals.size(); ++i) {
      const uint64_t b = i * width, wi = b / 64, off = b % 64;
      const uint64_t v = width == 64 ? vals[i] : (vals[i] & ((1ULL << width) - 1));
      w[wi] |= v << off;
      if (off + width > 64) w[wi + 1] |= v >> (64 - off);
    }
    u64(bits);
    const uint8_t wb = static_cast<uint8_t>(width);
    raw(&wb, 1);
    raw(w.data(), (bits + 63) / 64 * 8);
  }
  // sdsl rank_support_v5<1>: one int_vector<64>; per 2048-bit block a pair
  // (ones before the block, 12-bit prefix counts after every 6 words).
  void rank_v5(const vector<uint64_t> &w, uint64_t bits) {
    const uint64_t words = (bits + 63) / 64;
    const uint64_t entries = ((words * 64 >> 11) + 1) << 1;
    vector<uint64_t> bb(entries, 0);
    if (words) {
      uint64_t j = 0, sum = __builtin_popcountll(w[0]), second = 0, cnt = 1;
      for (uint64_t i = 1; i < words; ++i, ++cnt) {
        if (cnt == 32) {
          j += 2;
          bb[j - 1] = second;
          bb[j] = bb[j - 2] + sum;
          second = sum = cnt = 0;
        } else if (cnt % 6 == 0) {
          second |= sum << (60 - 12 * (cnt / 6));
        }
        sum += __builtin_popcountll(w[i]);
      }
      if (cnt % 6 == 0) second |= sum << (60 - 12 * (cnt / 6));
      if (cnt == 32) {
        j += 2;
        bb[j - 1] = second;
        bb[j] = bb[j - 2] + sum;
        if (j + 1 < entries) bb[j + 1] = 0;
      } else {
        bb[j + 1] = second;
      }
    }
    u64(entries * 64);
    raw(bb.data(), entries * 8);
  }
};

// Cumulative popcount per word: rank over [0, i) without a directory layout.
struct WordRank {
  vector<uint64_t> pre;
  const uint64_t *w = nullptr;
  void build(const vector<uint64_t> &bits) {
    w = bits.data();
    pre.assign(bits.size() + 1, 0);
    for (size_t i = 0; i < bits.size(); ++i) pre[i + 1] = pre[i] + __builtin_popcountll(bits[i]);
  }
  uint64_t rank(uint64_t i) const {
    uint64_t r = pre[i / 64];
    if (i % 64) r += __builtin_popcountll(w[i / 64] & ((1ULL << (i % 64)) - 1));
    return r;
  }
};

inline bool bit(const vector<uint64_t> &v, uint64_t i) { return (v[i / 64] >> (i % 64)) & 1; }
inline void set_bit(vector<uint64_t> &v, uint64_t i) { v[i / 64] |= 1ULL << (i % 64); }

// All k-mer keys (colex key: char j of the k-mer at bits [2j, 2j+1]) of the
// given sequences, forward and (optionally) reverse complement, in input order.
// Windows containing a code > 3 are skipped.
vector<uint64_t> kmer_keys(const uint8_t *codes, const uint64_t *ends, uint64_t n_seqs, uint32_t k,
                           bool add_rc, int T, vector<uint32_t> *seq_of = nullptr) {
  vector<uint64_t> offs(n_seqs + 1, 0);
  for (uint64_t s = 0; s < n_seqs; ++s) {
    const uint64_t st = s ? ends[s - 1] : 0, len = ends[s] - st;
    offs[s + 1] = offs[s] + (len >= k ? (len - k + 1) * (add_rc ? 2 : 1) : 0);
  }
  vector<uint64_t> keys(offs[n_seqs]);
  if (seq_of) seq_of->resize(offs[n_seqs]);
  const uint64_t mask = low_mask(k);
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
  for (uint64_t s = 0; s < n_seqs; ++s) {
    const uint64_t st = s ? ends[s - 1] : 0, len = ends[s] - st;
    if (len < k) continue;
    uint64_t o = offs[s], fw = 0, rc = 0, valid = 0;
    for (uint64_t i = 0; i < len; ++i) {
      const uint8_t c = codes[st + i];
      if (c > 3) {
        valid = 0;
        continue;
      }
      fw = (fw >> 2) | (static_cast<uint64_t>(c) << (2 * (k - 1)));
      rc = ((rc << 2) | (3 - c)) & mask;
      if (++valid >= k) {
        keys[o] = fw;
        if (seq_of) (*seq_of)[o] = static_cast<uint32_t>(s);
        ++o;
        if (add_rc) {
          keys[o] = rc;
          if (seq_of) (*seq_of)[o] = static_cast<uint32_t>(s);
          ++o;
        }
      }
    }
    // skipped windows hold UINT64_MAX, which sorts last and is dropped by callers
    for (uint64_t j = o; j < offs[s + 1]; ++j) {
      keys[j] = UINT64_MAX;
      if (seq_of) (*seq_of)[j] = static_cast<uint32_t>(s);
    }
  }
  return keys;
}

}  // namespace

// ===================================================================== RNG
extern "C" uint64_t bld_draw(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed * 0xD1B54A32D192ED03ULL + stream) + i);
}

extern "C" void bld_genome_random(uint64_t seed, uint64_t len, uint8_t *out) {
  const uint64_t words = (len + 31) / 32;
  const uint64_t base = mix64(seed * 0xD1B54A32D192ED03ULL + 0);
#pragma omp parallel for schedule(static)
  for (uint64_t w = 0; w < words; ++w) {
    const uint64_t x = mix64(base + w);
    for (uint64_t j = 0; j < 32 && w * 32 + j < len; ++j) out[w * 32 + j] = (x >> (2 * j)) & 3;
  }
}

extern "C" uint64_t bld_genome_mutate(const uint8_t *root, uint64_t len, uint64_t seed,
                                      double sub_rate, double indel_rate, uint32_t max_indel,
                                      uint8_t *out, uint64_t out_cap) {
  Rng r(seed, 1);
  uint64_t o = 0;
  for (uint64_t i = 0; i < len && o < out_cap;) {
    const double u = r.uniform();
    if (u < indel_rate && max_indel > 0) {
      const uint64_t l = 1 + r.next() % max_indel;
      if (r.next() & 1) {  // insertion after root[i]
        out[o++] = root[i++];
        for (uint64_t j = 0; j < l && o < out_cap; ++j) out[o++] = r.next() & 3;
      } else {  // deletion of l bases
        i += l;
      }
    } else if (u < indel_rate + sub_rate) {
      out[o++] = (root[i++] + 1 + r.next() % 3) & 3;
    } else {
      out[o++] = root[i++];
    }
  }
  return o;
}

extern "C" void bld_permutation(uint64_t seed, uint64_t n, uint64_t *out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = i;
  Rng r(seed, 9);
  for (uint64_t i = n; i > 1; --i) std::swap(out[i - 1], out[r.next() % i]);
}

extern "C" void bld_uniform_lengths(uint64_t seed, uint64_t n, uint32_t lo, uint32_t hi,
                                    uint32_t *out) {
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i) out[i] = lo + static_cast<uint32_t>(bld_draw(seed, 7, i) % (hi - lo + 1));
}

extern "C" void bld_lognormal_lengths(uint64_t seed, uint64_t n, double median, double sigma,
                                      uint32_t lo, uint32_t hi, uint32_t *out) {
  const double mu = std::log(median);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i) {
    double u1 = unit(bld_draw(seed, 8, 2 * i));
    const double u2 = unit(bld_draw(seed, 8, 2 * i + 1));
    if (u1 <= 0) u1 = 0x1.0p-53;
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    double v = std::exp(mu + sigma * z);
    v = std::min(std::max(v, static_cast<double>(lo)), static_cast<double>(hi));
    out[i] = static_cast<uint32_t>(std::llround(v));
  }
}

extern "C" uint64_t bld_reads(uint64_t seed, const uint8_t *genomes, const uint64_t *genome_ends,
                              uint64_t n_genomes, uint64_t n_reads, const uint32_t *lens,
                              double random_frac, double err_rate, double n_rate,
                              int mode_alternate, char *out, uint64_t *ends,
                              uint32_t *src_genome) {
  static const char kAcgt[4] = {'A', 'C', 'G', 'T'};
  uint64_t total = 0;
  for (uint64_t r = 0; r < n_reads; ++r) ends[r] = (total += lens[r]);
#pragma omp parallel for schedule(dynamic, 256)
  for (uint64_t r = 0; r < n_reads; ++r) {
    Rng g(seed, 1000 + r);
    const uint64_t len = lens[r], st = r ? ends[r - 1] : 0;
    char *o = out + st;
    bool random = mode_alternate ? (r % 2 == 1) : (g.uniform() < random_frac);
    uint64_t gi = 0, glen = 0, gst = 0;
    if (!random && n_genomes) {
      gi = g.next() % n_genomes;
      gst = gi ? genome_ends[gi - 1] : 0;
      glen = genome_ends[gi] - gst;
      if (glen < len) random = true;
    }
    if (random || !n_genomes) {
      for (uint64_t j = 0; j < len; ++j) o[j] = kAcgt[g.next() & 3];
      if (src_genome) src_genome[r] = UINT32_MAX;
    } else {
      const uint64_t start = g.next() % (glen - len + 1);
      const bool rc = g.next() & 1;
      const uint8_t *src = genomes + gst + start;
      for (uint64_t j = 0; j < len; ++j) {
        uint8_t b = rc ? static_cast<uint8_t>(3 - src[len - 1 - j]) : src[j];
        if (err_rate > 0 && g.uniform() < err_rate) b = (b + 1 + g.next() % 3) & 3;
        o[j] = kAcgt[b];
      }
      if (src_genome) src_genome[r] = static_cast<uint32_t>(gi);
    }
    if (n_rate > 0)
      for (uint64_t j = 0; j < len; ++j)
        if (g.uniform() < n_rate) o[j] = 'N';
  }
  return total;
}

// ===================================================================== SBWT
extern "C" bld_sbwt *bld_sbwt_build(const uint8_t *codes, const uint64_t *seq_ends, uint64_t n_seqs,
                                    uint32_t k, int add_rc, int threads) {
  if (k < 2 || k > 32) return nullptr;
  const int T = threads_or_default(threads);
  auto *S = new bld_sbwt();
  S->k = k;
  S->add_rc = add_rc != 0;
  // 1. sorted unique k-mers
  S->kmers = kmer_keys(codes, seq_ends, n_seqs, k, add_rc != 0, T);
  radix_sort(S->kmers.data(), S->kmers.size(), 64, T);
  S->kmers.erase(std::unique(S->kmers.begin(), S->kmers.end()), S->kmers.end());
  if (k < 32 && !S->kmers.empty() && S->kmers.back() == UINT64_MAX) S->kmers.pop_back();
  const vector<uint64_t> &K = S->kmers;
  const uint64_t nk = K.size();
  S->n_kmers = nk;
  // 2. (k-1)-prefixes with no k-mer having them as (k-1)-suffix
  const uint64_t pmask = low_mask(k - 1);
  vector<uint64_t> Q(nk);
#pragma omp parallel for num_threads(T)
  for (uint64_t i = 0; i < nk; ++i) Q[i] = K[i] & pmask;
  radix_sort(Q.data(), nk, 2 * (k - 1), T);
  Q.erase(std::unique(Q.begin(), Q.end()), Q.end());
  vector<uint64_t> dk;
  vector<uint8_t> dl;
  dk.push_back(0);  // the all-$ root is always node 0
  dl.push_back(0);
  for (uint64_t qi = 0, j = 0; qi < Q.size(); ++qi) {
    const uint64_t q = Q[qi];
    while (j < nk && (K[j] >> 2) < q) ++j;
    if (j < nk && (K[j] >> 2) == q) continue;
    for (uint32_t len = 1; len < k; ++len) {  // $^(k-len) p[0..len-1]
      dk.push_back((q & low_mask(len)) << (2 * (k - len)));
      dl.push_back(static_cast<uint8_t>(len));
    }
  }
  {
    vector<uint64_t> order(dk.size());
    for (uint64_t i = 0; i < dk.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
      return dk[a] != dk[b] ? dk[a] < dk[b] : dl[a] < dl[b];
    });
    for (uint64_t i = 0; i < order.size(); ++i) {
      const uint64_t o = order[i];
      if (i && dk[o] == S->dummy_key.back() && dl[o] == S->dummy_len.back()) continue;
      S->dummy_key.push_back(dk[o]);
      S->dummy_len.push_back(dl[o]);
    }
  }
  const uint64_t nd = S->dummy_key.size();
  const uint64_t n = nk + nd;
  S->n = n;
  // 3. merged node order
  vector<uint64_t> key(n);
  vector<uint8_t> len(n);
  {
    uint64_t i = 0, j = 0, o = 0;
    while (i < nk || j < nd) {
      const bool take_dummy = j < nd && (i >= nk || S->dummy_key[j] <= K[i]);
      if (take_dummy) {
        key[o] = S->dummy_key[j];
        len[o++] = S->dummy_len[j++];
      } else {
        key[o] = K[i++];
        len[o++] = static_cast<uint8_t>(k);
      }
    }
  }
  // 4. C map: node 0 is the root, then nodes by last char
  vector<uint64_t> cnt(4, 0);
  for (uint64_t v = 1; v < n; ++v) cnt[(key[v] >> (2 * (k - 1))) & 3]++;
  S->c_map[0] = 1;
  for (int c = 0; c < 4; ++c) S->c_map[c + 1] = S->c_map[c] + cnt[c];
  // 5. suffix groups: node string [1..k-1] as (primary >> 2, min(len, k-1))
  const uint64_t words = (n + 63) / 64;
  for (auto &b : S->bits) b.assign(words + 1, 0);
  vector<uint64_t> gstart;
  auto suf_len = [&](uint64_t v) { return std::min<uint64_t>(len[v], k - 1); };
  for (uint64_t v = 0; v < n; ++v) {
    if (v == 0 || (key[v] >> 2) != (key[v - 1] >> 2) || suf_len(v) != suf_len(v - 1)) {
      set_bit(S->bits[4], v);
      gstart.push_back(v);
    }
  }
  // 6. edges: node v (last char c) hangs off the group whose suffix is v[0..k-2]
  bool ok = true;
#pragma omp parallel for num_threads(4) reduction(&& : ok)
  for (int c = 0; c < 4; ++c) {
    uint64_t g = 0;
    for (uint64_t v = S->c_map[c]; v < S->c_map[c + 1]; ++v) {
      const uint64_t tk = key[v] & pmask, tl = len[v] - 1u;
      while (g < gstart.size()) {
        const uint64_t s = gstart[g];
        const uint64_t sk = key[s] >> 2, sl = suf_len(s);
        if (sk < tk || (sk == tk && sl < tl)) ++g;
        else break;
      }
      const uint64_t s = g < gstart.size() ? gstart[g] : 0;
      if (g >= gstart.size() || (key[s] >> 2) != tk || suf_len(s) != tl) {
        ok = false;
        break;
      }
      set_bit(S->bits[c], s);
    }
  }
  if (!ok) {
    delete S;
    return nullptr;
  }
  return S;
}

extern "C" void bld_sbwt_free(bld_sbwt *s) { delete s; }
extern "C" uint64_t bld_sbwt_num_nodes(const bld_sbwt *s) { return s->n; }
extern "C" uint64_t bld_sbwt_num_kmers(const bld_sbwt *s) { return s->n_kmers; }
extern "C" void bld_sbwt_bits(const bld_sbwt *s, int c, uint64_t *out) {
  std::memcpy(out, s->bits[c].data(), (s->n + 63) / 64 * 8);
}

extern "C" int bld_sbwt_lookup(const bld_sbwt *s, const uint64_t *keys, uint64_t n, uint64_t *out) {
  if (!s->has_kmers()) return -1;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i) {
    auto it = std::lower_bound(s->kmers.begin(), s->kmers.end(), keys[i]);
    out[i] = (it != s->kmers.end() && *it == keys[i]) ? s->node_of_kmer_index(it - s->kmers.begin())
                                                      : UINT64_MAX;
  }
  return 0;
}

// plain-matrix v0.1: the fields SbwtBuilder.cpp:58-104 reads or skips, in order.
extern "C" int bld_sbwt_write(const bld_sbwt *s, const char *path) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return -1;
  Out o{f};
  o.str("plain-matrix");
  o.str("v0.1");
  for (int c = 0; c < 4; ++c) o.bit_vector(s->bits[c], s->n);
  for (int c = 0; c < 4; ++c) o.rank_v5(s->bits[c], s->n);
  o.bit_vector(s->bits[4], s->n);
  o.u64(4 * 8);  // C array (sigma entries)
  for (int c = 0; c < 4; ++c) o.u64(s->c_map[c]);
  // precomputed intervals of all 8-mers, (-1,-1) if empty; 8-mer x has its first
  // char in the low bits (the order the reference's fixture index holds them in)
  const uint32_t pk = 8;
  WordRank wr[4];
  for (int c = 0; c < 4; ++c) wr[c].build(s->bits[c]);
  vector<int64_t> pre(2ULL << (2 * pk));
  for (uint64_t x = 0; x < (1ULL << (2 * pk)); ++x) {
    uint64_t l = 0, r = s->n - 1;
    bool empty = false;
    for (uint32_t sh = 0; sh < 2 * pk && !empty; sh += 2) {
      const int c = (x >> sh) & 3;
      l = s->c_map[c] + wr[c].rank(l);
      r = s->c_map[c] + wr[c].rank(r + 1) - 1;
      empty = l > r;
    }
    pre[2 * x] = empty ? -1 : static_cast<int64_t>(l);
    pre[2 * x + 1] = empty ? -1 : static_cast<int64_t>(r);
  }
  o.u64(pre.size() * 8);
  o.raw(pre.data(), pre.size() * 8);
  o.u64(pk);
  o.u64(s->n);
  o.u64(s->n_kmers);
  o.u64(s->k);
  const bool good = o.ok;
  return (std::fclose(f) == 0 && good) ? 0 : -1;
}

// ===================================================================== colors
extern "C" bld_colors *bld_colors_build(const bld_sbwt *s, const uint8_t *codes,
                                        const uint64_t *seq_ends, uint64_t n_seqs,
                                        const uint32_t *color_of, uint32_t d, int threads) {
  const int T = threads_or_default(threads);
  if (!s->has_kmers()) return nullptr;
  const uint32_t k = s->k;
  const uint64_t n = s->n;
  auto *C = new bld_colors();
  C->n = n;
  // 1. (k-mer, color) occurrences, sorted by k-mer then color
  vector<uint32_t> seq_of;
  vector<uint64_t> keys = kmer_keys(codes, seq_ends, n_seqs, k, s->add_rc, T, &seq_of);
  uint32_t max_color = 0;
  for (uint64_t g = 0; g < n_seqs; ++g) max_color = std::max(max_color, color_of[g]);
  C->num_colors = static_cast<uint64_t>(max_color) + 1;
  vector<uint32_t> col(keys.size());
#pragma omp parallel for num_threads(T)
  for (uint64_t i = 0; i < keys.size(); ++i) col[i] = color_of[seq_of[i]];
  seq_of.clear();
  seq_of.shrink_to_fit();
  {  // stable LSD: sort by color first, then by key
    vector<uint64_t> ck(keys.size());
    vector<uint32_t> idx(keys.size());
#pragma omp parallel for num_threads(T)
    for (uint64_t i = 0; i < keys.size(); ++i) {
      ck[i] = col[i];
      idx[i] = static_cast<uint32_t>(i);
    }
    if (keys.size() >= (1ULL << 32)) {
      delete C;
      return nullptr;
    }
    radix_sort_pairs(ck.data(), idx.data(), keys.size(), bits_for(max_color), T);
    vector<uint64_t> k2(keys.size());
#pragma omp parallel for num_threads(T)
    for (uint64_t i = 0; i < keys.size(); ++i) {
      k2[i] = keys[idx[i]];
      idx[i] = col[idx[i]];
    }
    keys.swap(k2);
    radix_sort_pairs(keys.data(), idx.data(), keys.size(), 64, T);
    col.swap(idx);
  }
  // 2. color set of every real node, deduplicated into set ids. Pairs are
  // sorted by (k-mer, color); drop repeated pairs, then the distinct k-mers of
  // the pair list are exactly s->kmers in order.
  {
    uint64_t w = 0;
    for (uint64_t i = 0; i < keys.size(); ++i) {
      if (keys[i] == UINT64_MAX && k < 32) break;  // skipped windows sort last
      if (w && keys[w - 1] == keys[i] && col[w - 1] == col[i]) continue;
      keys[w] = keys[i];
      col[w] = col[i];
      ++w;
    }
    keys.resize(w);
    col.resize(w);
  }
  const uint64_t nk = s->n_kmers;
  vector<uint64_t> run(nk + 1, 0);  // colors of k-mer i: col[run[i] .. run[i+1])
  {
    uint64_t r = 0;
    for (uint64_t j = 0; j < keys.size(); ++j)
      if (j == 0 || keys[j] != keys[j - 1]) run[r++] = j;
    if (r != nk) {
      delete C;
      return nullptr;
    }
    run[nk] = keys.size();
  }
  keys.clear();
  keys.shrink_to_fit();
  vector<uint64_t> h(nk);
  vector<uint32_t> order(nk);
#pragma omp parallel for schedule(static) num_threads(T)
  for (uint64_t i = 0; i < nk; ++i) {
    uint64_t x = 0x12345678ULL + (run[i + 1] - run[i]);
    for (uint64_t j = run[i]; j < run[i + 1]; ++j) x = mix64(x ^ col[j]);
    h[i] = x;
    order[i] = static_cast<uint32_t>(i);
  }
  if (nk >= (1ULL << 32)) {
    delete C;
    return nullptr;
  }
  radix_sort_pairs(h.data(), order.data(), nk, 64, T);
  vector<uint32_t> kmer_set(nk);
  C->set_offsets.push_back(0);
  auto same_list = [&](uint64_t a, uint64_t b) {
    const uint64_t la = run[a + 1] - run[a], lb = run[b + 1] - run[b];
    return la == lb && std::equal(col.begin() + run[a], col.begin() + run[a + 1], col.begin() + run[b]);
  };
  for (uint64_t i = 0; i < nk;) {
    uint64_t e = i;
    while (e < nk && h[e] == h[i]) ++e;
    // equal hashes: one set per distinct list (lists compared exactly)
    vector<uint64_t> heads;
    vector<uint32_t> ids;
    for (uint64_t q = i; q < e; ++q) {
      const uint64_t kq = order[q];
      uint32_t id = UINT32_MAX;
      for (size_t t = 0; t < heads.size(); ++t)
        if (same_list(heads[t], kq)) id = ids[t];
      if (id == UINT32_MAX) {
        id = static_cast<uint32_t>(C->set_offsets.size() - 1);
        C->set_ids.insert(C->set_ids.end(), col.begin() + run[kq], col.begin() + run[kq + 1]);
        C->set_offsets.push_back(C->set_ids.size());
        heads.push_back(kq);
        ids.push_back(id);
      }
      kmer_set[kq] = id;
    }
    i = e;
  }
  h.clear();
  h.shrink_to_fit();
  order.clear();
  order.shrink_to_fit();
  col.clear();
  col.shrink_to_fit();
  run.clear();
  run.shrink_to_fit();
  vector<uint32_t> node_set(n, UINT32_MAX);
#pragma omp parallel for schedule(static) num_threads(T)
  for (uint64_t i = 0; i < nk; ++i) node_set[s->node_of_kmer_index(i)] = kmer_set[i];
  kmer_set.clear();
  kmer_set.shrink_to_fit();
  // 3. key k-mer marks
  const uint64_t words = (n + 63) / 64;
  C->key_marks.assign(words + 1, 0);
  auto mark = [&](uint64_t v) {
    __atomic_fetch_or(&C->key_marks[v >> 6], 1ULL << (v & 63), __ATOMIC_RELAXED);
  };
  auto marked = [&](uint64_t v) {
    return (__atomic_load_n(&C->key_marks[v >> 6], __ATOMIC_RELAXED) >> (v & 63)) & 1;
  };
  WordRank wr[4];
  for (int c = 0; c < 4; ++c) wr[c].build(s->bits[c]);
  vector<uint64_t> succ(n, UINT64_MAX);  // unique successor of unmarked real nodes
#pragma omp parallel for schedule(static, 4096) num_threads(T)
  for (uint64_t v = 0; v < n; ++v) {
    if (node_set[v] == UINT32_MAX) continue;  // dummy
    const bool alone = bit(s->bits[4], v) && (v + 1 == n || bit(s->bits[4], v + 1));
    int outdeg = 0, c1 = 0;
    for (int c = 0; c < 4; ++c)
      if (bit(s->bits[c], v)) {
        ++outdeg;
        c1 = c;
      }
    if (!alone || outdeg != 1) {
      mark(v);
      continue;
    }
    const uint64_t w = s->c_map[c1] + wr[c1].rank(v);
    if (node_set[w] != node_set[v]) {
      mark(v);
      continue;
    }
    succ[v] = w;
  }
  // every d-th node along the remaining chains (in-degree <= 1, so chains are
  // disjoint and walk in parallel), then the cycles
  vector<uint8_t> has_pred(n, 0), seen(n, 0);
#pragma omp parallel for schedule(static, 4096) num_threads(T)
  for (uint64_t v = 0; v < n; ++v)
    if (succ[v] != UINT64_MAX && !marked(succ[v])) has_pred[succ[v]] = 1;
  const uint32_t dd = d ? d : 1;
  auto walk = [&](uint64_t v, bool cycle) {
    uint64_t steps = 0;
    if (cycle) {
      mark(v);
      seen[v] = 1;
      v = succ[v];
    }
    while (v != UINT64_MAX && !seen[v] && !marked(v)) {
      seen[v] = 1;
      if (++steps % dd == 0) mark(v);
      v = succ[v];
    }
  };
#pragma omp parallel for schedule(dynamic, 4096) num_threads(T)
  for (uint64_t v = 0; v < n; ++v)
    if (succ[v] != UINT64_MAX && !has_pred[v] && !marked(v)) walk(v, false);
  for (uint64_t v = 0; v < n; ++v)
    if (succ[v] != UINT64_MAX && !marked(v) && !seen[v]) walk(v, true);
  succ.clear();
  succ.shrink_to_fit();
  // 4. color_set_idxs over key nodes; dense/sparse choice per set
  for (uint64_t v = 0; v < n; ++v)
    if (bit(C->key_marks, v)) C->color_set_idxs.push_back(node_set[v] == UINT32_MAX ? 0 : node_set[v]);
  sbwt_build::choose_dense(C);
  return C;
}

extern "C" void bld_colors_free(bld_colors *c) { delete c; }
extern "C" uint64_t bld_colors_num_colors(const bld_colors *c) { return c->num_colors; }
extern "C" uint64_t bld_colors_num_sets(const bld_colors *c) { return c->set_offsets.size() - 1; }
extern "C" uint64_t bld_colors_num_dense(const bld_colors *c) { return c->n_dense; }
extern "C" void bld_colors_key_marks(const bld_colors *c, uint64_t *out) {
  std::memcpy(out, c->key_marks.data(), (c->n + 63) / 64 * 8);
}

// sdsl-hybrid-v4: the fields ColorIndexBuilder.cpp:23-86 reads, in order.
extern "C" int bld_colors_write(const bld_colors *c, const char *path) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return -1;
  Out o{f};
  const uint64_t nsets = c->set_offsets.size() - 1;
  vector<uint64_t> dense_bits, dense_iv{0}, sparse_vals, sparse_iv{0};
  uint64_t dense_len = 0;
  vector<uint64_t> is_dense((nsets + 63) / 64 + 1, 0);
  for (uint64_t i = 0; i < nsets; ++i) {
    const uint64_t a = c->set_offsets[i], b = c->set_offsets[i + 1];
    if (c->dense[i]) {
      set_bit(is_dense, i);
      const uint64_t width = c->set_ids[b - 1] + 1;  // bitmap up to the largest id
      dense_bits.resize((dense_len + width + 63) / 64 + 1, 0);
      for (uint64_t j = a; j < b; ++j) set_bit(dense_bits, dense_len + c->set_ids[j]);
      dense_len += width;
      dense_iv.push_back(dense_len);
    } else {
      for (uint64_t j = a; j < b; ++j) sparse_vals.push_back(c->set_ids[j]);
      sparse_iv.push_back(sparse_vals.size());
    }
  }
  dense_bits.resize((dense_len + 63) / 64 + 1, 0);
  o.str("sdsl-hybrid-v4");
  o.bit_vector(dense_bits, dense_len);
  o.int_vector(dense_iv, bits_for(dense_len));
  o.int_vector(sparse_vals, bits_for(c->num_colors - 1));
  o.int_vector(sparse_iv, bits_for(sparse_vals.size()));
  o.bit_vector(is_dense, nsets);
  o.rank_v5(is_dense, nsets);
  o.bit_vector(c->key_marks, c->n);
  o.rank_v5(c->key_marks, c->n);
  o.int_vector(c->color_set_idxs, bits_for(nsets ? nsets - 1 : 0));
  o.u64(nsets ? nsets - 1 : 0);
  o.u64(c->num_colors - 1);
  o.u64(c->set_ids.size());
  const bool good = o.ok;
  return (std::fclose(f) == 0 && good) ? 0 : -1;
}
