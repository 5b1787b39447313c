// GPU construction of the SBWT (plain-matrix) and the color index
// (sdsl-hybrid-v4) on one MI355X: the same nodes, bit vectors, color sets, set
// ids and key k-mer marks as the host builder in builder.cpp, computed from a
// device sort of all k-mer occurrences. Sized for genome collections of
// billions of bases: 1000 genomes of 2-5 Mbp hold about 7 G k-mer occurrences,
// whose (key, color) pairs sort in HBM.
//
// Steps (device unless noted):
//  1. k-mer keys (forward and reverse complement) of every window, one 16-byte
//     pair per position, with the color of the window's sequence;
//  2. stable radix sort of (key, color), sortedness checked;
//  3. distinct keys (the real k-mers) and distinct (key, color) pairs with the
//     per-k-mer runs of color ids;
//  4. (k-1)-prefixes that are no k-mer's (k-1)-suffix, found through a
//     24-bit bucket table over the sorted suffixes; their dummies sorted on the
//     host (they are a few per sequence);
//  5. suffix-group starts (one wave-word per 64 nodes) and the edge bits, each
//     real k-mer setting its letter on its source group's first node;
//  6. color-set ids: a hash per color list, deduplicated in a device hash table,
//     ids in hash order (the host builder's order), lists compared exactly;
//  7. key k-mer marks: branching nodes and color-set changes by one ballot per
//     64 nodes, then every d-th node along each remaining chain, one lane per
//     chain, and the rare cycles on the host in node order.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "builder.h"
#include "builder_impl.hpp"

using std::vector;
using u64 = uint64_t;
using u32 = uint32_t;
using u8 = uint8_t;

namespace {

constexpr u64 kMax = ~0ULL;
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr u32 kTag = 0x80000000u;  // first color of a k-mer's run (colors are below 2^31)
constexpr u32 kBlock = 256;

#define GB_HIP(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));              \
  } while (0)

__host__ __device__ inline u64 mix64(u64 z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ inline u64 mask_digits(u32 digits) {
  return digits >= 32 ? ~0ULL : ((1ULL << (2 * digits)) - 1);
}

// Grid-stride loops throughout: a launch's grid is counted in work-items in a
// 32-bit field, so one lane per element would cut off launches beyond 2^32.
__device__ inline u64 lane0() { return static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ inline u64 lanes() { return static_cast<u64>(gridDim.x) * blockDim.x; }

inline u32 grid_for(u64 threads) {
  const u64 g = (threads + kBlock - 1) / kBlock;
  return static_cast<u32>(std::min<u64>(std::max<u64>(g, 1), 1u << 20));
}

// Device memory with a running total, so the build reports its peak.
struct Arena {
  u64 cur = 0, peak = 0;
};

template <class T>
struct DBuf {
  T *p = nullptr;
  u64 n = 0;
  Arena *a = nullptr;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  void alloc(Arena &ar, u64 count) {
    release();
    a = &ar;
    n = count;
    const u64 bytes = std::max<u64>(count, 1) * sizeof(T);
    GB_HIP(hipMalloc(&p, bytes));
    a->cur += bytes;
    a->peak = std::max(a->peak, a->cur);
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      a->cur -= std::max<u64>(n, 1) * sizeof(T);
    }
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

struct Temp {  // rocPRIM temporary storage, grown on demand
  DBuf<u8> b;
  Arena *a;
  explicit Temp(Arena &ar) : a(&ar) {}
  void *get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > b.n) b.alloc(*a, bytes);
    return b.p;
  }
};

// ---------------------------------------------------------------- 1. k-mer keys
constexpr u32 kKeyChunk = 32;  // windows per lane

// Every position q of the concatenated sequences ends one window; its keys go
// to slots 2q (forward) and 2q+1 (reverse complement), or slot q without
// reverse complements. Windows that are shorter than k or hold a code > 3
// get kMax (k <= 31, so no k-mer key is kMax).
__global__ __launch_bounds__(kBlock) void k_kmer_keys(const u8 *__restrict__ codes, u64 L,
                                                      const u64 *__restrict__ ends, u64 n_seqs,
                                                      const u32 *__restrict__ color_of, u32 k,
                                                      int add_rc, u64 *__restrict__ keys,
                                                      u32 *__restrict__ cols) {
  for (u64 p0 = lane0() * kKeyChunk; p0 < L; p0 += lanes() * kKeyChunk) {
  const u64 p1 = p0 + kKeyChunk < L ? p0 + kKeyChunk : L;
  u64 lo = 0, hi = n_seqs;  // first sequence ending after p0
  while (lo < hi) {
    const u64 m = (lo + hi) >> 1;
    if (ends[m] > p0) hi = m;
    else lo = m + 1;
  }
  u64 s = lo;
  const u64 st = s ? ends[s - 1] : 0;
  u64 q = p0 >= st + (k - 1) ? p0 - (k - 1) : st;
  const u64 mask = mask_digits(k);
  u64 fw = 0, rc = 0;
  u32 valid = 0;
  for (; q < p1; ++q) {
    while (q >= ends[s]) {  // a new sequence starts at q
      ++s;
      fw = rc = 0;
      valid = 0;
    }
    const u32 c = codes[q];
    if (c > 3) {
      valid = 0;
    } else {
      fw = (fw >> 2) | (static_cast<u64>(c) << (2 * (k - 1)));
      rc = ((rc << 2) | (3 - c)) & mask;
      ++valid;
    }
    if (q >= p0) {
      const bool ok = valid >= k;
      const u32 col = color_of ? color_of[s] : 0;
      if (add_rc) {
        keys[2 * q] = ok ? fw : kMax;
        keys[2 * q + 1] = ok ? rc : kMax;
        if (cols) {
          cols[2 * q] = col;
          cols[2 * q + 1] = col;
        }
      } else {
        keys[q] = ok ? fw : kMax;
        if (cols) cols[q] = col;
      }
    }
  }
  }
}

// Sorted by key, and by color within a key when colors ride along.
__global__ __launch_bounds__(kBlock) void k_check_sorted(const u64 *__restrict__ keys,
                                                         const u32 *__restrict__ cols, u64 n,
                                                         u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i >= n) return;
    const u64 a = keys[i - 1], b = keys[i];
    if (a > b || (cols && a == b && cols[i - 1] > cols[i])) atomicOr(err, 1u);
  };
  for (u64 t = lane0() + 1; t < n; t += lanes()) body(t);
}

// Multiset fingerprints: the wrapping sum of a hash of every element (of the
// keys, the (key, color) pairs or the colors; of flagged elements only when
// flags are given). A sort or a compaction that loses or repeats an element
// changes it.
__device__ inline u64 element_print(const u64 *keys, const u32 *cols, u64 i) {
  if (!keys) return mix64((cols[i] & ~kTag) + 0x51ED27ULL);
  if (!cols) return mix64(keys[i]);
  return mix64(keys[i] ^ mix64((cols[i] & ~kTag) + 0x51ED27ULL));
}

__global__ __launch_bounds__(kBlock) void k_fingerprint(const u64 *__restrict__ keys,
                                                        const u32 *__restrict__ cols,
                                                        const u8 *__restrict__ flags, u64 n,
                                                        unsigned long long *__restrict__ out) {
  u64 acc = 0;
  for (u64 i = lane0(); i < n; i += lanes())
    if (!flags || flags[i]) acc += element_print(keys, cols, i);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(static_cast<unsigned long long>(acc), o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, static_cast<unsigned long long>(acc));
}

// Strictly increasing keys (the distinct k-mers).
__global__ __launch_bounds__(kBlock) void k_check_increasing(const u64 *__restrict__ keys, u64 n,
                                                             u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i < n && keys[i - 1] >= keys[i]) atomicOr(err, 1u);
  };
  for (u64 t = lane0() + 1; t < n; t += lanes()) body(t);
}

// flag_k: first occurrence of a real k-mer key. With colors, flag_p: first
// occurrence of a (key, color) pair, and the pair's color tagged when it is
// its k-mer's first.
__global__ __launch_bounds__(kBlock) void k_distinct_flags(const u64 *__restrict__ keys,
                                                           u32 *__restrict__ cols, u64 n,
                                                           u8 *__restrict__ flag_k,
                                                           u8 *__restrict__ flag_p) {
  const auto body = [&](const u64 i) {
    if (i >= n) return;
    const u64 key = keys[i];
    const bool valid = key != kMax;
    const bool new_key = valid && (i == 0 || keys[i - 1] != key);
    if (flag_k) flag_k[i] = new_key;
    if (flag_p) {
      const u32 c = cols[i] & ~kTag;
      const bool new_pair = valid && (new_key || (cols[i - 1] & ~kTag) != c);
      flag_p[i] = new_pair;
      if (new_key) cols[i] = c | kTag;
    }
  };
  for (u64 t = lane0(); t < n; t += lanes()) body(t);
}

struct TagFlag {
  __host__ __device__ u8 operator()(u32 c) const { return (c & kTag) ? 1 : 0; }
};

__global__ __launch_bounds__(kBlock) void k_strip_tags(u32 *__restrict__ cols, u64 n) {
  const auto body = [&](const u64 i) {
    if (i < n) cols[i] &= ~kTag;
  };
  for (u64 t = lane0(); t < n; t += lanes()) body(t);
}

// ---------------------------------------------------------------- 4-5. SBWT
struct SuffixTable {  // bucket b = top `bits` bits of a (k-1)-mer
  const u64 *K;
  const u64 *T;  // T[b] = first k-mer whose suffix (K >> 2) falls in bucket >= b
  u64 nk;
  u32 shift;
  __device__ u64 lower(u64 q) const {  // first i with (K[i] >> 2) >= q
    const u64 b = q >> shift;
    u64 lo = T[b], hi = T[b + 1];
    while (lo < hi) {
      const u64 m = (lo + hi) >> 1;
      if ((K[m] >> 2) < q) lo = m + 1;
      else hi = m;
    }
    return lo;
  }
};

__global__ __launch_bounds__(kBlock) void k_bucket_table(const u64 *__restrict__ K, u64 nk,
                                                         u32 shift, u64 buckets,
                                                         u64 *__restrict__ T) {
  const auto body = [&](const u64 b) {
    if (b > buckets) return;
    u64 lo = 0, hi = nk;
    while (lo < hi) {
      const u64 m = (lo + hi) >> 1;
      if (((K[m] >> 2) >> shift) < b) lo = m + 1;
      else hi = m;
    }
    T[b] = lo;
  };
  for (u64 t = lane0(); t < buckets + 1; t += lanes()) body(t);
}

// (k-1)-prefixes of k-mers that are no k-mer's (k-1)-suffix (with repeats).
__global__ __launch_bounds__(kBlock) void k_missing_prefixes(SuffixTable st, u64 pmask,
                                                             u64 *__restrict__ out, u64 cap,
                                                             unsigned long long *__restrict__ count) {
  const auto body = [&](const u64 i) {
    if (i >= st.nk) return;
    const u64 q = st.K[i] & pmask;
    const u64 j = st.lower(q);
    if (j < st.nk && (st.K[j] >> 2) == q) return;
    const u64 slot = atomicAdd(count, 1ULL);
    if (slot < cap) out[slot] = q;
  };
  for (u64 t = lane0(); t < st.nk; t += lanes()) body(t);
}

__device__ inline u64 upper_u64(const u64 *a, u64 n, u64 x) {  // # a[j] <= x
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 m = (lo + hi) >> 1;
    if (a[m] <= x) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__device__ inline u64 lower_u64(const u64 *a, u64 n, u64 x) {  // # a[j] < x
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 m = (lo + hi) >> 1;
    if (a[m] < x) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_real_before(const u64 *__restrict__ K, u64 nk,
                                                        const u64 *__restrict__ dkey, u64 nd,
                                                        u64 *__restrict__ out) {
  const auto body = [&](const u64 j) {
    if (j < nd) out[j] = lower_u64(K, nk, dkey[j]);
  };
  for (u64 t = lane0(); t < nd; t += lanes()) body(t);
}

struct Nodes {  // the merged node order: real k-mers and dummies
  const u64 *K;
  u64 nk;
  const u64 *dkey;
  const u8 *dlen;
  const u64 *dpos;  // node index of each dummy, ascending
  u64 nd, n;
  u32 k;
};

// Suffix-group starts: one lane per 64 nodes writes one word. A node's group
// key is (primary >> 2, min(len, k - 1)); a start differs from its predecessor.
__global__ __launch_bounds__(kBlock) void k_group_starts(Nodes N, u64 words,
                                                         u64 *__restrict__ sgs) {
  const auto body = [&](const u64 w) {
    if (w >= words) return;
    const u64 v0 = w * 64;
    u64 t = lower_u64(N.dpos, N.nd, v0);  // dummies before v0
    u64 pk = kMax;
    u32 pl = 0xFFFFFFFFu;
    if (v0 > 0) {  // the node before v0
      if (t > 0 && N.dpos[t - 1] == v0 - 1) {
        pk = N.dkey[t - 1] >> 2;
        pl = min(static_cast<u32>(N.dlen[t - 1]), N.k - 1);
      } else {
        pk = N.K[v0 - 1 - t] >> 2;
        pl = N.k - 1;
      }
    }
    u64 word = 0;
    const u64 v1 = v0 + 64 < N.n ? v0 + 64 : N.n;
    for (u64 v = v0; v < v1; ++v) {
      u64 ck;
      u32 cl;
      if (t < N.nd && N.dpos[t] == v) {
        ck = N.dkey[t] >> 2;
        cl = min(static_cast<u32>(N.dlen[t]), N.k - 1);
        ++t;
      } else {
        ck = N.K[v - t] >> 2;
        cl = N.k - 1;
      }
      if (v == 0 || ck != pk || cl != pl) word |= 1ULL << (v & 63);
      pk = ck;
      pl = cl;
    }
    sgs[w] = word;
  };
  for (u64 t = lane0(); t < words; t += lanes()) body(t);
}

// Edges into real k-mers: k-mer v = x[0..k-1] hangs off the first node of the
// group whose suffix is x[0..k-2] (a run of real k-mers, or the dummy
// $x[0..k-2] when no k-mer ends with it) and sets letter x[k-1] there.
__global__ __launch_bounds__(kBlock) void k_real_edges(SuffixTable st, Nodes N, u64 pmask,
                                                       u64 *__restrict__ b0, u64 *__restrict__ b1,
                                                       u64 *__restrict__ b2, u64 *__restrict__ b3,
                                                       u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i >= N.nk) return;
    const u64 key = N.K[i];
    const u64 q = key & pmask;
    const u32 c = static_cast<u32>(key >> (2 * (N.k - 1))) & 3;
    const u64 j = st.lower(q);
    u64 s;
    if (j < N.nk && (N.K[j] >> 2) == q) {
      s = j + upper_u64(N.dkey, N.nd, N.K[j]);
    } else {  // the dummy (q << 2, k - 1)
      const u64 dk = q << 2;
      u64 lo = 0, hi = N.nd;
      while (lo < hi) {
        const u64 m = (lo + hi) >> 1;
        if (N.dkey[m] < dk || (N.dkey[m] == dk && N.dlen[m] < N.k - 1)) lo = m + 1;
        else hi = m;
      }
      if (lo >= N.nd || N.dkey[lo] != dk || N.dlen[lo] != N.k - 1) {
        atomicOr(err, 2u);
        return;
      }
      s = N.dpos[lo];
    }
    u64 *b = c == 0 ? b0 : c == 1 ? b1 : c == 2 ? b2 : b3;
    atomicOr(reinterpret_cast<unsigned long long *>(b + (s >> 6)), 1ULL << (s & 63));
  };
  for (u64 t = lane0(); t < N.nk; t += lanes()) body(t);
}

__global__ __launch_bounds__(kBlock) void k_set_bits(const u64 *__restrict__ list, u64 n,
                                                     u64 *__restrict__ b0, u64 *__restrict__ b1,
                                                     u64 *__restrict__ b2, u64 *__restrict__ b3) {
  const auto body = [&](const u64 i) {
    if (i >= n) return;
    const u64 e = list[i];  // node << 2 | letter
    const u64 s = e >> 2;
    const u32 c = e & 3;
    u64 *b = c == 0 ? b0 : c == 1 ? b1 : c == 2 ? b2 : b3;
    atomicOr(reinterpret_cast<unsigned long long *>(b + (s >> 6)), 1ULL << (s & 63));
  };
  for (u64 t = lane0(); t < n; t += lanes()) body(t);
}

// ---------------------------------------------------------------- 6. color sets
__global__ __launch_bounds__(kBlock) void k_list_hash(const u64 *__restrict__ run,
                                                      const u32 *__restrict__ cols, u64 nk,
                                                      u64 *__restrict__ h, u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i >= nk) return;
    const u64 a = run[i], b = run[i + 1];
    u64 x = 0x12345678ULL + (b - a);
    for (u64 j = a; j < b; ++j) x = mix64(x ^ cols[j]);
    if (x == kMax) atomicOr(err, 4u);  // kMax marks an empty table slot
    h[i] = x;
  };
  for (u64 t = lane0(); t < nk; t += lanes()) body(t);
}

struct HashTable {
  u64 *keys;  // kMax = empty
  u64 *reps;  // smallest k-mer index holding the hash
  u32 *ids;   // set id of the slot
  u64 mask;
  __device__ u64 slot_of(u64 h) const {
    u64 s = h & mask;
    for (u64 probes = 0; probes <= mask; ++probes) {
      const u64 cur = keys[s];
      if (cur == h) return s;
      if (cur == kMax) return kMax;
      s = (s + 1) & mask;
    }
    return kMax;
  }
};

__global__ __launch_bounds__(kBlock) void k_table_insert(HashTable t, const u64 *__restrict__ h,
                                                         u64 nk,
                                                         unsigned long long *__restrict__ distinct,
                                                         u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i >= nk) return;
    const u64 x = h[i];
    u64 s = x & t.mask;
    for (u64 probes = 0; probes <= t.mask; ++probes) {
      const unsigned long long prev =
          atomicCAS(reinterpret_cast<unsigned long long *>(t.keys + s), kMax, x);
      if (prev == kMax || prev == x) {
        if (prev == kMax) atomicAdd(distinct, 1ULL);
        atomicMin(reinterpret_cast<unsigned long long *>(t.reps + s), i);
        return;
      }
      s = (s + 1) & t.mask;
    }
    atomicOr(err, 8u);  // table full
  };
  for (u64 t = lane0(); t < nk; t += lanes()) body(t);
}

struct Occupied {
  __host__ __device__ u8 operator()(u64 key) const { return key != kMax; }
};

__global__ __launch_bounds__(kBlock) void k_table_ids(HashTable t, const u64 *__restrict__ hs,
                                                      u64 m) {
  const auto body = [&](const u64 id) {
    if (id >= m) return;
    const u64 s = t.slot_of(hs[id]);
    if (s != kMax) t.ids[s] = static_cast<u32>(id);
  };
  for (u64 t = lane0(); t < m; t += lanes()) body(t);
}

__global__ __launch_bounds__(kBlock) void k_set_sizes(const u64 *__restrict__ reps,
                                                      const u64 *__restrict__ run, u64 m,
                                                      u64 *__restrict__ sizes) {
  const auto body = [&](const u64 id) {
    if (id < m) sizes[id] = run[reps[id] + 1] - run[reps[id]];
  };
  for (u64 t = lane0(); t < m; t += lanes()) body(t);
}

__global__ __launch_bounds__(kBlock) void k_set_lists(const u64 *__restrict__ reps,
                                                      const u64 *__restrict__ run,
                                                      const u64 *__restrict__ offs,
                                                      const u32 *__restrict__ cols, u64 m,
                                                      u32 *__restrict__ ids) {
  const auto body = [&](const u64 id) {
    if (id >= m) return;
    const u64 a = run[reps[id]], b = run[reps[id] + 1];
    u64 o = offs[id];
    for (u64 j = a; j < b; ++j) ids[o++] = cols[j];
  };
  for (u64 t = lane0(); t < m; t += lanes()) body(t);
}

// node_set[node of k-mer i] = its set id; the k-mer's list must equal the
// list of the set's first k-mer (two lists with one hash are not supported).
__global__ __launch_bounds__(kBlock) void k_node_sets(HashTable t, const u64 *__restrict__ h,
                                                      const u64 *__restrict__ run,
                                                      const u32 *__restrict__ cols, u64 nk,
                                                      const u64 *__restrict__ real_before, u64 nd,
                                                      u32 *__restrict__ node_set,
                                                      u32 *__restrict__ err) {
  const auto body = [&](const u64 i) {
    if (i >= nk) return;
    const u64 s = t.slot_of(h[i]);
    if (s == kMax) {
      atomicOr(err, 16u);
      return;
    }
    const u64 rep = t.reps[s];
    if (rep != i) {
      const u64 a = run[i], b = run[i + 1], ra = run[rep], rb = run[rep + 1];
      bool same = b - a == rb - ra;
      for (u64 j = 0; same && j < b - a; ++j) same = cols[a + j] == cols[ra + j];
      if (!same) atomicOr(err, 32u);
    }
    node_set[i + upper_u64(real_before, nd, i)] = t.ids[s];
  };
  for (u64 t = lane0(); t < nk; t += lanes()) body(t);
}

// ---------------------------------------------------------------- 7. key marks
struct Graph {
  const u64 *bits[5];  // A, C, G, T, group starts
  const u64 *pre[4];   // ones before each word
  const u32 *node_set;
  u64 c_map[4];
  u64 n;
  __device__ bool bit(int c, u64 v) const { return (bits[c][v >> 6] >> (v & 63)) & 1; }
  __device__ u64 rank(int c, u64 v) const {
    const u64 w = v >> 6, r = v & 63;
    return pre[c][w] + (r ? __popcll(bits[c][w] & ((1ULL << r) - 1)) : 0);
  }
  // The forward successor of a real node that is alone in its group and has
  // one out-edge, else kMax.
  __device__ u64 single_successor(u64 v) const {
    const bool alone = bit(4, v) && (v + 1 == n || bit(4, v + 1));
    if (!alone) return kMax;
    int outdeg = 0, c1 = 0;
    for (int c = 0; c < 4; ++c)
      if (bit(c, v)) {
        ++outdeg;
        c1 = c;
      }
    if (outdeg != 1) return kMax;
    return c_map[c1] + rank(c1, v);
  }
};

__device__ inline bool test_bit(const u64 *b, u64 v) { return (b[v >> 6] >> (v & 63)) & 1; }
__device__ inline void or_bit(u64 *b, u64 v) {
  atomicOr(reinterpret_cast<unsigned long long *>(b + (v >> 6)), 1ULL << (v & 63));
}

__global__ __launch_bounds__(kBlock) void k_popcounts(const u64 *__restrict__ b, u64 words,
                                                      u64 *__restrict__ out) {
  const auto body = [&](const u64 w) {
    if (w < words) out[w] = __popcll(b[w]);
  };
  for (u64 t = lane0(); t < words; t += lanes()) body(t);
}

// Marks of nodes that end a chain by themselves: a real node that shares its
// group, has out-degree != 1, or precedes a node of another color set. One
// ballot per 64 nodes (the wave covers one word).
__global__ __launch_bounds__(kBlock) void k_marks_local(Graph g, u64 words,
                                                        u64 *__restrict__ marks) {
  for (u64 v = lane0(); v < words * 64; v += lanes()) {  // whole waves: the stride is 64-aligned
    bool m = false;
    if (v < g.n && g.node_set[v] != kNone) {
      const u64 w = g.single_successor(v);
      m = w == kMax || g.node_set[w] != g.node_set[v];
    }
    const u64 word = __ballot(m);
    if ((threadIdx.x & 63) == 0) marks[v >> 6] = word;
  }
}

__device__ inline bool chain_node(const Graph &g, const u64 *marks1, u64 v) {
  return g.node_set[v] != kNone && !test_bit(marks1, v);
}

__global__ __launch_bounds__(kBlock) void k_has_pred(Graph g, const u64 *__restrict__ marks1,
                                                     u64 *__restrict__ has_pred) {
  const auto body = [&](const u64 v) {
    if (v >= g.n || !chain_node(g, marks1, v)) return;
    const u64 w = g.single_successor(v);
    if (!test_bit(marks1, w)) or_bit(has_pred, w);
  };
  for (u64 t = lane0(); t < g.n; t += lanes()) body(t);
}

// One lane per chain: from a chain's first node, mark every d-th node until a
// node that is marked by itself.
__global__ __launch_bounds__(kBlock) void k_walk_chains(Graph g, const u64 *__restrict__ marks1,
                                                        const u64 *__restrict__ has_pred, u32 d,
                                                        u64 *__restrict__ marks,
                                                        u64 *__restrict__ seen,
                                                        u32 *__restrict__ err) {
  const auto body = [&](const u64 v) {
    if (v >= g.n || !chain_node(g, marks1, v) || test_bit(has_pred, v)) return;
    u64 u = v, steps = 0;
    while (!test_bit(marks1, u)) {
      if (steps >= g.n) {  // a chain holds at most n nodes
        atomicOr(err, 64u);
        return;
      }
      or_bit(seen, u);
      if (++steps % d == 0) or_bit(marks, u);
      u = g.single_successor(u);
    }
  };
  for (u64 t = lane0(); t < g.n; t += lanes()) body(t);
}

__global__ __launch_bounds__(kBlock) void k_cycle_nodes(Graph g, const u64 *__restrict__ marks1,
                                                        const u64 *__restrict__ seen,
                                                        u64 *__restrict__ out, u64 cap,
                                                        unsigned long long *__restrict__ count) {
  const auto body = [&](const u64 v) {
    if (v >= g.n || !chain_node(g, marks1, v) || test_bit(seen, v)) return;
    const u64 slot = atomicAdd(count, 1ULL);
    if (slot < cap) {
      out[2 * slot] = v;
      out[2 * slot + 1] = g.single_successor(v);
    }
  };
  for (u64 t = lane0(); t < g.n; t += lanes()) body(t);
}

__global__ __launch_bounds__(kBlock) void k_mark_list(const u64 *__restrict__ list, u64 n,
                                                      u64 *__restrict__ marks) {
  const auto body = [&](const u64 i) {
    if (i < n) or_bit(marks, list[i]);
  };
  for (u64 t = lane0(); t < n; t += lanes()) body(t);
}

struct MarkFlag {
  const u64 *marks;
  __host__ __device__ u8 operator()(u64 v) const { return (marks[v >> 6] >> (v & 63)) & 1; }
};

struct SetIdx {
  __host__ __device__ u32 operator()(u32 s) const { return s == kNone ? 0 : s; }
};

// ---------------------------------------------------------------- host driver
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

u32 read_err(const DBuf<u32> &e) {
  u32 v = 0;
  GB_HIP(hipMemcpy(&v, e.p, 4, hipMemcpyDeviceToHost));
  return v;
}

u64 read_u64(const u64 *p) {
  u64 v = 0;
  GB_HIP(hipMemcpy(&v, p, 8, hipMemcpyDeviceToHost));
  return v;
}

u64 fingerprint(const u64 *keys, const u32 *cols, const u8 *flags, u64 n, DBuf<u64> &acc) {
  GB_HIP(hipMemset(acc.p, 0, 8));
  if (n) {
    k_fingerprint<<<static_cast<u32>(std::min<u64>((n + kBlock - 1) / kBlock, 8192)), kBlock>>>(
        keys, cols, flags, n, reinterpret_cast<unsigned long long *>(acc.p));
    GB_HIP(hipGetLastError());
  }
  return read_u64(acc.p);
}

template <class In, class Flags, class Out>
u64 select_flagged(Temp &tmp, In in, Flags flags, Out out, u64 n, DBuf<u64> &count) {
  if (n == 0) return 0;
  size_t bytes = 0;
  GB_HIP(rocprim::select(nullptr, bytes, in, flags, out, count.p, n));
  GB_HIP(rocprim::select(tmp.get(bytes), bytes, in, flags, out, count.p, n));
  return read_u64(count.p);
}

template <class In>
void exclusive_scan_u64(Temp &tmp, In in, u64 *out, u64 n) {
  if (n == 0) return;
  size_t bytes = 0;
  GB_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, u64{0}, n, rocprim::plus<u64>()));
  GB_HIP(rocprim::exclusive_scan(tmp.get(bytes), bytes, in, out, u64{0}, n,
                                 rocprim::plus<u64>()));
}

// Sequences in color order (stable), so a stable sort by key leaves each
// k-mer's colors ascending.
void color_ordered(const u8 *codes, const u64 *ends, u64 n_seqs, const u32 *color_of,
                   vector<u8> &c2, vector<u64> &e2, vector<u32> &col2) {
  vector<u64> ord(n_seqs);
  for (u64 s = 0; s < n_seqs; ++s) ord[s] = s;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](u64 a, u64 b) { return color_of[a] < color_of[b]; });
  c2.resize(n_seqs ? ends[n_seqs - 1] : 0);
  e2.resize(n_seqs);
  col2.resize(n_seqs);
  u64 o = 0;
  for (u64 i = 0; i < n_seqs; ++i) {
    const u64 s = ord[i], st = s ? ends[s - 1] : 0, len = ends[s] - st;
    std::memcpy(c2.data() + o, codes + st, len);
    o += len;
    e2[i] = o;
    col2[i] = color_of[s];
  }
}

struct Built {
  bld_sbwt *S = nullptr;
  bld_colors *C = nullptr;
};

thread_local const char *g_stage = "start";  // named in error messages

Built build(const u8 *codes_in, const u64 *ends_in, u64 n_seqs, u32 k, bool add_rc,
            const u32 *color_of_in, u32 d, bool keep_kmers, bld_gpu_stats *stats) {
  if (k < 2 || k > 31) throw std::invalid_argument("GPU build supports k in [2, 31]");
  const auto t_all = Clock::now();
  bld_gpu_stats dummy_stats{};
  bld_gpu_stats &st = stats ? *stats : dummy_stats;
  st = bld_gpu_stats{};
  Arena ar;
  Temp tmp(ar);
  const bool colors = color_of_in != nullptr;
  // sequences in color order when the colors are not already ascending
  vector<u8> c2;
  vector<u64> e2;
  vector<u32> col2;
  const u8 *codes = codes_in;
  const u64 *ends = ends_in;
  const u32 *color_of = color_of_in;
  u32 max_color = 0;
  if (colors) {
    bool sorted = true;
    for (u64 s = 0; s < n_seqs; ++s) {
      if (color_of[s] >= kTag) throw std::invalid_argument("GPU build needs color ids below 2^31");
      max_color = std::max(max_color, color_of[s]);
      if (s && color_of[s] < color_of[s - 1]) sorted = false;
    }
    if (!sorted) {
      color_ordered(codes_in, ends_in, n_seqs, color_of_in, c2, e2, col2);
      codes = c2.data();
      ends = e2.data();
      color_of = col2.data();
    }
  }
  const u64 L = n_seqs ? ends[n_seqs - 1] : 0;
  const u64 N = L * (add_rc ? 2 : 1);
  st.occurrences = N;

  g_stage = "keys";
  auto t0 = Clock::now();
  DBuf<u64> count;
  count.alloc(ar, 1);
  DBuf<u32> err;
  err.alloc(ar, 1);
  GB_HIP(hipMemset(err.p, 0, 4));
  // ---- 1-2. keys and sort
  DBuf<u64> ka, kb;
  DBuf<u32> ca, cb;
  {
    DBuf<u8> d_codes;
    DBuf<u64> d_ends;
    DBuf<u32> d_col;
    d_codes.alloc(ar, L);
    d_ends.alloc(ar, n_seqs);
    if (L) GB_HIP(hipMemcpy(d_codes.p, codes, L, hipMemcpyHostToDevice));
    if (n_seqs) GB_HIP(hipMemcpy(d_ends.p, ends, n_seqs * 8, hipMemcpyHostToDevice));
    if (colors) {
      d_col.alloc(ar, n_seqs);
      if (n_seqs) GB_HIP(hipMemcpy(d_col.p, color_of, n_seqs * 4, hipMemcpyHostToDevice));
    }
    ka.alloc(ar, N);
    kb.alloc(ar, N);
    if (colors) {
      ca.alloc(ar, N);
      cb.alloc(ar, N);
    }
    if (L) {
      k_kmer_keys<<<grid_for((L + kKeyChunk - 1) / kKeyChunk), kBlock>>>(
          d_codes.p, L, d_ends.p, n_seqs, colors ? d_col.p : nullptr, k, add_rc ? 1 : 0, ka.p,
          colors ? ca.p : nullptr);
      GB_HIP(hipGetLastError());
    }
    GB_HIP(hipDeviceSynchronize());
    st.ms_keys = ms_since(t0);
  }
  g_stage = "sort";
  t0 = Clock::now();
  const u64 print_in = fingerprint(ka.p, colors ? ca.p : nullptr, nullptr, N, count);
  if (N > 1) {
    rocprim::double_buffer<u64> kbuf(ka.p, kb.p);
    size_t bytes = 0;
    if (colors) {
      rocprim::double_buffer<u32> cbuf(ca.p, cb.p);
      GB_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kbuf, cbuf, N, 0, 64));
      GB_HIP(rocprim::radix_sort_pairs(tmp.get(bytes), bytes, kbuf, cbuf, N, 0, 64));
      if (cbuf.current() != ca.p) std::swap(ca.p, cb.p);
    } else {
      GB_HIP(rocprim::radix_sort_keys(nullptr, bytes, kbuf, N, 0, 64));
      GB_HIP(rocprim::radix_sort_keys(tmp.get(bytes), bytes, kbuf, N, 0, 64));
    }
    if (kbuf.current() != ka.p) std::swap(ka.p, kb.p);  // sorted data in ka / ca
    k_check_sorted<<<grid_for(N), kBlock>>>(ka.p, colors ? ca.p : nullptr, N, err.p);
    GB_HIP(hipGetLastError());
    if (read_err(err)) throw std::runtime_error("device sort produced an unsorted sequence");
  }
  if (fingerprint(ka.p, colors ? ca.p : nullptr, nullptr, N, count) != print_in)
    throw std::runtime_error("device sort lost or repeated elements");
  GB_HIP(hipDeviceSynchronize());
  st.ms_sort = ms_since(t0);
  t0 = Clock::now();
  g_stage = "distinct";
  // ---- 3. distinct k-mers (into kb) and (k-mer, color) pairs (into cb)
  u64 nk = 0, P = 0;
  DBuf<u64> run;
  {
    DBuf<u8> fk, fp;
    fk.alloc(ar, N);
    if (colors) fp.alloc(ar, N);
    if (N) {
      k_distinct_flags<<<grid_for(N), kBlock>>>(ka.p, colors ? ca.p : nullptr, N, fk.p,
                                                colors ? fp.p : nullptr);
      GB_HIP(hipGetLastError());
    }
    nk = select_flagged(tmp, ka.p, fk.p, kb.p, N, count);
    if (fingerprint(ka.p, nullptr, fk.p, N, count) != fingerprint(kb.p, nullptr, nullptr, nk, count))
      throw std::runtime_error("distinct k-mer compaction lost or repeated elements");
    if (nk > 1) {
      k_check_increasing<<<grid_for(nk), kBlock>>>(kb.p, nk, err.p);
      GB_HIP(hipGetLastError());
      if (read_err(err)) throw std::runtime_error("distinct k-mers are not strictly increasing");
    }
    if (colors) {
      P = select_flagged(tmp, ca.p, fp.p, cb.p, N, count);
      if (fingerprint(nullptr, ca.p, fp.p, N, count) !=
          fingerprint(nullptr, cb.p, nullptr, P, count))
        throw std::runtime_error("distinct pair compaction lost or repeated elements");
    }
  }
  ka.release();
  ca.release();
  if (colors) {
    run.alloc(ar, nk + 1);
    const u64 runs = select_flagged(tmp, rocprim::counting_iterator<u64>(0),
                                    rocprim::make_transform_iterator(cb.p, TagFlag()), run.p, P,
                                    count);
    if (runs != nk) throw std::runtime_error("k-mer runs do not match the distinct k-mers");
    GB_HIP(hipMemcpy(run.p + nk, &P, 8, hipMemcpyHostToDevice));
    if (P) {
      k_strip_tags<<<grid_for(P), kBlock>>>(cb.p, P);
      GB_HIP(hipGetLastError());
    }
  }
  const u64 *K = kb.p;
  st.kmers = nk;
  st.pairs = P;

  g_stage = "dummies";
  // ---- 4. dummies
  auto *S = new bld_sbwt();
  std::unique_ptr<bld_sbwt> S_guard(S);
  S->k = k;
  S->add_rc = add_rc;
  S->n_kmers = nk;
  const u64 pmask = mask_digits(k - 1);
  const u32 tbits = std::min<u32>(24, 2 * (k - 1));
  const u32 shift = 2 * (k - 1) - tbits;
  const u64 buckets = 1ULL << tbits;
  DBuf<u64> T;
  T.alloc(ar, buckets + 1);
  k_bucket_table<<<grid_for(buckets + 1), kBlock>>>(K, nk, shift, buckets, T.p);
  GB_HIP(hipGetLastError());
  const SuffixTable sufs{K, T.p, nk, shift};
  vector<u64> missing;
  {
    u64 cap = std::min<u64>(std::max<u64>(nk, 1), 1ULL << 24);
    for (;;) {
      DBuf<u64> out;
      out.alloc(ar, cap);
      GB_HIP(hipMemset(count.p, 0, 8));
      if (nk) {
        k_missing_prefixes<<<grid_for(nk), kBlock>>>(sufs, pmask, out.p, cap,
                                                     reinterpret_cast<unsigned long long *>(count.p));
        GB_HIP(hipGetLastError());
      }
      const u64 m = read_u64(count.p);
      if (m <= cap) {
        missing.resize(m);
        if (m) GB_HIP(hipMemcpy(missing.data(), out.p, m * 8, hipMemcpyDeviceToHost));
        break;
      }
      cap = m;
    }
  }
  std::sort(missing.begin(), missing.end());
  missing.erase(std::unique(missing.begin(), missing.end()), missing.end());
  st.missing_prefixes = missing.size();
  {
    vector<std::pair<u64, u8>> dm;
    dm.reserve(1 + missing.size() * (k - 1));
    dm.emplace_back(0, 0);  // the all-$ root is always node 0
    for (u64 q : missing)
      for (u32 len = 1; len < k; ++len)  // $^(k-len) p[0..len-1]
        dm.emplace_back((q & mask_digits(len)) << (2 * (k - len)), static_cast<u8>(len));
    std::sort(dm.begin(), dm.end());
    dm.erase(std::unique(dm.begin(), dm.end()), dm.end());
    S->dummy_key.resize(dm.size());
    S->dummy_len.resize(dm.size());
    for (size_t j = 0; j < dm.size(); ++j) {
      S->dummy_key[j] = dm[j].first;
      S->dummy_len[j] = dm[j].second;
    }
  }
  const u64 nd = S->dummy_key.size();
  const u64 n = nk + nd;
  S->n = n;
  st.dummies = nd;
  st.nodes = n;
  DBuf<u64> dkey, dpos, real_before;
  DBuf<u8> dlen;
  dkey.alloc(ar, nd);
  dlen.alloc(ar, nd);
  dpos.alloc(ar, nd);
  real_before.alloc(ar, nd);
  GB_HIP(hipMemcpy(dkey.p, S->dummy_key.data(), nd * 8, hipMemcpyHostToDevice));
  GB_HIP(hipMemcpy(dlen.p, S->dummy_len.data(), nd, hipMemcpyHostToDevice));
  k_real_before<<<grid_for(nd), kBlock>>>(K, nk, dkey.p, nd, real_before.p);
  GB_HIP(hipGetLastError());
  vector<u64> rb(nd), pos(nd);
  GB_HIP(hipMemcpy(rb.data(), real_before.p, nd * 8, hipMemcpyDeviceToHost));
  for (u64 j = 0; j < nd; ++j) pos[j] = j + rb[j];
  GB_HIP(hipMemcpy(dpos.p, pos.data(), nd * 8, hipMemcpyHostToDevice));

  g_stage = "edges";
  // ---- 5. C map, group starts, edges
  {
    u64 cnt[4] = {0, 0, 0, 0};
    vector<u64> bound(5, 0);
    bound[4] = nk;
    if (nk) {
      vector<u64> probe = {1ULL << (2 * (k - 1)), 2ULL << (2 * (k - 1)), 3ULL << (2 * (k - 1))};
      DBuf<u64> pk, po;
      pk.alloc(ar, 3);
      po.alloc(ar, 3);
      GB_HIP(hipMemcpy(pk.p, probe.data(), 24, hipMemcpyHostToDevice));
      k_real_before<<<1, kBlock>>>(K, nk, pk.p, 3, po.p);
      GB_HIP(hipGetLastError());
      GB_HIP(hipMemcpy(bound.data() + 1, po.p, 24, hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < 4; ++c) cnt[c] = bound[c + 1] - bound[c];
    for (u64 j = 1; j < nd; ++j) cnt[(S->dummy_key[j] >> (2 * (k - 1))) & 3]++;
    S->c_map[0] = 1;
    for (int c = 0; c < 4; ++c) S->c_map[c + 1] = S->c_map[c] + cnt[c];
  }
  const u64 words = (n + 63) / 64;
  DBuf<u64> bits[5];
  for (auto &b : bits) {
    b.alloc(ar, words + 1);
    GB_HIP(hipMemset(b.p, 0, (words + 1) * 8));
  }
  const Nodes nodes{K, nk, dkey.p, dlen.p, dpos.p, nd, n, k};
  k_group_starts<<<grid_for(words), kBlock>>>(nodes, words, bits[4].p);
  GB_HIP(hipGetLastError());
  if (nk) {
    k_real_edges<<<grid_for(nk), kBlock>>>(sufs, nodes, pmask, bits[0].p, bits[1].p, bits[2].p,
                                          bits[3].p, err.p);
    GB_HIP(hipGetLastError());
  }
  {  // edges into dummies come from dummies: the host holds them all
    std::map<std::pair<u64, u32>, u64> group_of;  // (suffix, length) -> first node
    for (u64 j = 0; j < nd; ++j) {
      const std::pair<u64, u32> key{S->dummy_key[j] >> 2,
                                    std::min<u32>(S->dummy_len[j], k - 1)};
      group_of.emplace(key, pos[j]);  // keeps the first
    }
    vector<u64> list;
    for (u64 j = 1; j < nd; ++j) {
      const u32 c = (S->dummy_key[j] >> (2 * (k - 1))) & 3;
      auto it = group_of.find({S->dummy_key[j] & pmask, static_cast<u32>(S->dummy_len[j]) - 1});
      if (it == group_of.end()) throw std::runtime_error("a dummy node has no source group");
      list.push_back(it->second << 2 | c);
    }
    if (!list.empty()) {
      DBuf<u64> dl;
      dl.alloc(ar, list.size());
      GB_HIP(hipMemcpy(dl.p, list.data(), list.size() * 8, hipMemcpyHostToDevice));
      k_set_bits<<<grid_for(list.size()), kBlock>>>(dl.p, list.size(), bits[0].p, bits[1].p,
                                                    bits[2].p, bits[3].p);
      GB_HIP(hipGetLastError());
    }
  }
  if (read_err(err)) throw std::runtime_error("a k-mer has no source group");
  for (int c = 0; c < 5; ++c) {
    S->bits[c].assign(words + 1, 0);
    GB_HIP(hipMemcpy(S->bits[c].data(), bits[c].p, words * 8, hipMemcpyDeviceToHost));
  }
  if (keep_kmers) {
    S->kmers.resize(nk);
    if (nk) GB_HIP(hipMemcpy(S->kmers.data(), K, nk * 8, hipMemcpyDeviceToHost));
  }
  GB_HIP(hipDeviceSynchronize());
  st.ms_sbwt = ms_since(t0);
  T.release();
  kb.release();  // the k-mer keys are not needed past here
  Built out;
  if (colors) {
    t0 = Clock::now();
    auto *C = new bld_colors();
    std::unique_ptr<bld_colors> C_guard(C);
    C->n = n;
    C->num_colors = static_cast<u64>(max_color) + 1;
  g_stage = "color sets";
    // ---- 6. color sets
    DBuf<u64> h;
    h.alloc(ar, nk);
    if (nk) {
      k_list_hash<<<grid_for(nk), kBlock>>>(run.p, cb.p, nk, h.p, err.p);
      GB_HIP(hipGetLastError());
    }
    if (read_err(err)) throw std::runtime_error("a color list hashes to the reserved value");
    DBuf<u64> tk, tr;
    DBuf<u32> tid;
    u64 cap = 1ULL << 16, m = 0;
    while (cap < nk / 4) cap <<= 1;
    for (;;) {
      tk.alloc(ar, cap);
      tr.alloc(ar, cap);
      GB_HIP(hipMemset(tk.p, 0xFF, cap * 8));
      GB_HIP(hipMemset(tr.p, 0xFF, cap * 8));
      GB_HIP(hipMemset(count.p, 0, 8));
      HashTable t{tk.p, tr.p, nullptr, cap - 1};
      if (nk) {
        k_table_insert<<<grid_for(nk), kBlock>>>(t, h.p, nk,
                                                 reinterpret_cast<unsigned long long *>(count.p),
                                                 err.p);
        GB_HIP(hipGetLastError());
      }
      m = read_u64(count.p);
      const u32 e = read_err(err);
      if (!(e & 8u) && m <= cap / 10 * 8) break;
      GB_HIP(hipMemset(err.p, 0, 4));
      cap <<= 1;
    }
    st.sets = m;
    tid.alloc(ar, cap);
    HashTable table{tk.p, tr.p, tid.p, cap - 1};
    DBuf<u64> hs, rs, offs;
    DBuf<u32> set_ids;
    {
      DBuf<u64> hs2, rs2;
      hs.alloc(ar, m);
      rs.alloc(ar, m);
      hs2.alloc(ar, m);
      rs2.alloc(ar, m);
      auto occ = rocprim::make_transform_iterator(tk.p, Occupied());
      if (select_flagged(tmp, tk.p, occ, hs.p, cap, count) != m ||
          select_flagged(tmp, tr.p, occ, rs.p, cap, count) != m)
        throw std::runtime_error("hash table compaction lost entries");
      if (m > 1) {  // set ids in hash order
        rocprim::double_buffer<u64> kbuf(hs.p, hs2.p), vbuf(rs.p, rs2.p);
        size_t bytes = 0;
        GB_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kbuf, vbuf, m, 0, 64));
        GB_HIP(rocprim::radix_sort_pairs(tmp.get(bytes), bytes, kbuf, vbuf, m, 0, 64));
        if (kbuf.current() != hs.p) std::swap(hs.p, hs2.p);
        if (vbuf.current() != rs.p) std::swap(rs.p, rs2.p);
      }
    }
    if (m) {
      k_table_ids<<<grid_for(m), kBlock>>>(table, hs.p, m);
      GB_HIP(hipGetLastError());
    }
    hs.release();
    offs.alloc(ar, m + 1);
    {
      DBuf<u64> sizes;
      sizes.alloc(ar, m + 1);
      GB_HIP(hipMemset(sizes.p, 0, (m + 1) * 8));
      if (m) {
        k_set_sizes<<<grid_for(m), kBlock>>>(rs.p, run.p, m, sizes.p);
        GB_HIP(hipGetLastError());
      }
      exclusive_scan_u64(tmp, sizes.p, offs.p, m + 1);
    }
    const u64 total_ids = read_u64(offs.p + m);
    set_ids.alloc(ar, total_ids);
    if (m) {
      k_set_lists<<<grid_for(m), kBlock>>>(rs.p, run.p, offs.p, cb.p, m, set_ids.p);
      GB_HIP(hipGetLastError());
    }
    C->set_offsets.resize(m + 1);
    C->set_ids.resize(total_ids);
    GB_HIP(hipMemcpy(C->set_offsets.data(), offs.p, (m + 1) * 8, hipMemcpyDeviceToHost));
    if (total_ids)
      GB_HIP(hipMemcpy(C->set_ids.data(), set_ids.p, total_ids * 4, hipMemcpyDeviceToHost));
    set_ids.release();
    offs.release();
    rs.release();
    DBuf<u32> node_set;
    node_set.alloc(ar, n);
    GB_HIP(hipMemset(node_set.p, 0xFF, n * 4));
    if (nk) {
      k_node_sets<<<grid_for(nk), kBlock>>>(table, h.p, run.p, cb.p, nk, real_before.p, nd,
                                           node_set.p, err.p);
      GB_HIP(hipGetLastError());
    }
    const u32 e = read_err(err);
    if (e & 32u) throw std::runtime_error("two distinct color lists share a hash value");
    if (e) throw std::runtime_error("a color list is missing from the hash table");
    h.release();
    tk.release();
    tr.release();
    tid.release();
    run.release();
    cb.release();
  g_stage = "key marks";
    // ---- 7. key k-mer marks
    DBuf<u64> pre[4];
    Graph g{};
    for (int c = 0; c < 4; ++c) {
      pre[c].alloc(ar, words + 1);
      DBuf<u64> pc;
      pc.alloc(ar, words + 1);
      GB_HIP(hipMemset(pc.p, 0, (words + 1) * 8));
      k_popcounts<<<grid_for(words), kBlock>>>(bits[c].p, words, pc.p);
      GB_HIP(hipGetLastError());
      exclusive_scan_u64(tmp, pc.p, pre[c].p, words + 1);
      g.bits[c] = bits[c].p;
      g.pre[c] = pre[c].p;
      g.c_map[c] = S->c_map[c];
    }
    g.bits[4] = bits[4].p;
    g.node_set = node_set.p;
    g.n = n;
    DBuf<u64> marks1, marks, has_pred, seen;
    for (auto *b : {&marks1, &marks, &has_pred, &seen}) {
      b->alloc(ar, words + 1);
      GB_HIP(hipMemset(b->p, 0, (words + 1) * 8));
    }
    k_marks_local<<<grid_for(words * 64), kBlock>>>(g, words, marks1.p);
    GB_HIP(hipGetLastError());
    GB_HIP(hipMemcpy(marks.p, marks1.p, (words + 1) * 8, hipMemcpyDeviceToDevice));
    k_has_pred<<<grid_for(n), kBlock>>>(g, marks1.p, has_pred.p);
    GB_HIP(hipGetLastError());
    const u32 dd = d ? d : 1;
    k_walk_chains<<<grid_for(n), kBlock>>>(g, marks1.p, has_pred.p, dd, marks.p, seen.p, err.p);
    GB_HIP(hipGetLastError());
    if (read_err(err)) throw std::runtime_error("a chain of the successor graph does not end");
    {  // cycles: every node of a cycle has a predecessor, so no chain walk reached it
      vector<u64> cyc;
      u64 ccap = 1ULL << 20;
      for (;;) {
        DBuf<u64> out;
        out.alloc(ar, 2 * ccap);
        GB_HIP(hipMemset(count.p, 0, 8));
        k_cycle_nodes<<<grid_for(n), kBlock>>>(g, marks1.p, seen.p, out.p, ccap,
                                               reinterpret_cast<unsigned long long *>(count.p));
        GB_HIP(hipGetLastError());
        const u64 c = read_u64(count.p);
        if (c <= ccap) {
          cyc.resize(2 * c);
          if (c) GB_HIP(hipMemcpy(cyc.data(), out.p, 2 * c * 8, hipMemcpyDeviceToHost));
          break;
        }
        ccap = c;
      }
      st.cycle_nodes = cyc.size() / 2;
      if (!cyc.empty()) {  // the host builder's order: cycles by their smallest node
        std::unordered_map<u64, u64> succ;
        vector<u64> vs;
        for (size_t i = 0; i < cyc.size(); i += 2) {
          succ[cyc[i]] = cyc[i + 1];
          vs.push_back(cyc[i]);
        }
        std::sort(vs.begin(), vs.end());
        std::unordered_map<u64, bool> seen_h, marked_h;
        vector<u64> new_marks;
        for (u64 v : vs) {
          if (marked_h.count(v) || seen_h.count(v)) continue;
          marked_h[v] = true;
          new_marks.push_back(v);
          seen_h[v] = true;
          u64 u = succ.at(v), steps = 0;
          while (!seen_h.count(u) && !marked_h.count(u)) {
            seen_h[u] = true;
            if (++steps % dd == 0) {
              marked_h[u] = true;
              new_marks.push_back(u);
            }
            u = succ.at(u);
          }
        }
        DBuf<u64> ml;
        ml.alloc(ar, new_marks.size());
        GB_HIP(hipMemcpy(ml.p, new_marks.data(), new_marks.size() * 8, hipMemcpyHostToDevice));
        k_mark_list<<<grid_for(new_marks.size()), kBlock>>>(ml.p, new_marks.size(), marks.p);
        GB_HIP(hipGetLastError());
      }
    }
    C->key_marks.assign(words + 1, 0);
    GB_HIP(hipMemcpy(C->key_marks.data(), marks.p, words * 8, hipMemcpyDeviceToHost));
    {  // set id of every key node, in node order
      u64 key_nodes = 0;
      for (u64 w = 0; w < words; ++w) key_nodes += __builtin_popcountll(C->key_marks[w]);
      DBuf<u32> idx;
      idx.alloc(ar, key_nodes);
      const u64 got = select_flagged(
          tmp, rocprim::make_transform_iterator(node_set.p, SetIdx()),
          rocprim::make_transform_iterator(rocprim::counting_iterator<u64>(0), MarkFlag{marks.p}),
          idx.p, n, count);
      if (got != key_nodes) throw std::runtime_error("key node count mismatch");
      vector<u32> h_idx(key_nodes);
      if (key_nodes) GB_HIP(hipMemcpy(h_idx.data(), idx.p, key_nodes * 4, hipMemcpyDeviceToHost));
      C->color_set_idxs.assign(h_idx.begin(), h_idx.end());
      st.key_nodes = key_nodes;
    }
    sbwt_build::choose_dense(C);
    GB_HIP(hipDeviceSynchronize());
    st.ms_colors = ms_since(t0);
    out.C = C_guard.release();
  }
  st.peak_device_bytes = ar.peak;
  st.ms_total = ms_since(t_all);
  out.S = S_guard.release();
  return out;
}

}  // namespace

extern "C" bld_sbwt *bld_gpu_build(const uint8_t *codes, const uint64_t *seq_ends,
                                   uint64_t n_seqs, uint32_t k, int add_rc,
                                   const uint32_t *color_of, uint32_t d, int keep_kmers,
                                   bld_colors **colors, bld_gpu_stats *stats, char *err,
                                   uint64_t err_len) {
  if (colors) *colors = nullptr;
  g_stage = "start";
  try {
    if (color_of && !colors) throw std::invalid_argument("colors requested without an output");
    Built b = build(codes, seq_ends, n_seqs, k, add_rc != 0, color_of, d, keep_kmers != 0, stats);
    if (colors) *colors = b.C;
    return b.S;
  } catch (const std::exception &e) {
    if (err && err_len) std::snprintf(err, err_len, "%s (stage: %s)", e.what(), g_stage);
    return nullptr;
  }
}
