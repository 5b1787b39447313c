// Device index construction: ACGT rank blocks and mark entries (built on the
// host, uploaded once) and the presearch table (built on the device).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "device_fns.hpp"
#include "device_index.hpp"
#include "loaders.hpp"

namespace sbwt_amd {

void throw_hip(hipError_t e, const char *file, int line) {
  throw std::runtime_error(std::string("GPUassert: ") + hipGetErrorString(e) + " at " + file + ":" +
                           std::to_string(line));
}

u32 default_presearch_letters(u32 k, u64 n, bool wide, u64 device_bytes) {
  const u32 lo = std::min<u32>(k, 12), hi = std::min<u32>(k, 16);
  u32 log4 = 0;
  while (log4 < 32 && (1ULL << (2 * log4)) < n) ++log4;  // ceil(log4(n))
  u32 kp = std::clamp<u32>(log4 + 2, lo, hi);
  const u64 entry = wide ? 16 : 8;
  while (kp > lo && (entry << (2 * kp)) > device_bytes / 8) --kp;
  return kp;
}

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

// Key-k-mer move destination of every node (the walk of
// src/IndexSearcher/IndexSearcher.cuh:60-70 done once per node), so a found
// k-mer's move costs one table read. Walks longer than kKeyTableSteps are
// left to the query (kKeyLong*).
template <typename Idx>
__global__ __launch_bounds__(256) void k_key_dest(IndexView ix, Idx *out, u64 n, u32 cap) {
  for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 d = key_walk_capped<Idx>(ix, static_cast<Idx>(i), cap);
    if constexpr (sizeof(Idx) == 4) {
      out[i] = d == kKeyLong64 ? kKeyLong32 : (d == kWalkStuck ? kKeyStuck32 : static_cast<u32>(d));
    } else {
      out[i] = d;
    }
  }
}

// ---- k-mer table: the non-empty intervals of every prefix length from kp
// up to k, level by level (each level has at most one entry per node: the
// intervals of distinct strings of one length are disjoint), then the k-letter
// intervals (single nodes) inserted by code.
struct PrefixEnt {
  u64 code;
  u32 l, r;
};

// Reserves cnt output slots for each thread of the workgroup, in thread
// order, with one atomic per workgroup (a single counter takes ~10^8 atomics
// per second, so per-wave appends would bound the build). Every thread of the
// workgroup calls it.
__device__ __forceinline__ u64 block_reserve(u32 cnt, u64 *n_out) {
  __shared__ u32 wave_sums[4];
  __shared__ u64 block_base;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32 incl = cnt;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u32 total = wave_sums[0] + wave_sums[1] + wave_sums[2] + wave_sums[3];
    block_base = total ? atomicAdd(reinterpret_cast<unsigned long long *>(n_out), static_cast<u64>(total)) : 0;
  }
  __syncthreads();
  u32 before = 0;
  for (u32 w = 0; w < wave; ++w) before += wave_sums[w];
  const u64 base = block_base + before + incl - cnt;
  __syncthreads();  // the shared words are reused by the next call
  return base;
}

constexpr u32 kSeedItems = 16;  // presearch entries per thread

// The non-empty presearch intervals of entries [first, first + count)
// (nothing is written past cap; the count goes on).
__global__ __launch_bounds__(256) void k_hash_seed(const uint2 *__restrict__ pre, u64 first,
                                                   u64 count, PrefixEnt *out, u64 *n_out, u64 cap) {
  pre += first;
  const u64 tile = 256ULL * kSeedItems;
  for (u64 b = static_cast<u64>(blockIdx.x) * tile; b < count; b += static_cast<u64>(gridDim.x) * tile) {
    const u64 x0 = b + threadIdx.x * kSeedItems;
    u32 keep = 0;
    for (u32 q = 0; q < kSeedItems; ++q)
      if (x0 + q < count) {
        const uint2 e = pre[x0 + q];
        keep += e.x <= e.y ? 1u : 0u;
      }
    u64 at = block_reserve(keep, n_out);
    for (u32 q = 0; q < kSeedItems && keep; ++q)
      if (x0 + q < count) {
        const uint2 e = pre[x0 + q];
        if (e.x <= e.y) {
          if (at < cap) out[at] = PrefixEnt{first + x0 + q, e.x, e.y};
          ++at;
          --keep;
        }
      }
  }
}

// One letter more for every entry of the level; at the last level (k
// letters) the single-node intervals go into the table.
__global__ __launch_bounds__(256) void k_hash_extend(IndexView ix, const PrefixEnt *__restrict__ in,
                                                     const u64 *n_in, PrefixEnt *out, u64 *n_out,
                                                     u64 cap, u32 last, uint4 *table, u32 bits,
                                                     u32 node_bits) {
  const u64 n = min(*n_in, cap);
  const u64 gmask = (1ULL << bits) - 1;
  const u32 rem_bits = 2 * ix.k - bits;
  for (u64 b = static_cast<u64>(blockIdx.x) * 256; b < n; b += static_cast<u64>(gridDim.x) * 256) {
    const u64 i = b + threadIdx.x;
    PrefixEnt e{0, 1, 0};
    if (i < n) e = in[i];
    PrefixEnt ch[4];
    u32 keep = 0;
#pragma unroll
    for (u32 c = 0; c < 4; ++c) {
      u32 l = e.l, r = e.r;
      if (l <= r) sbwt_step<u32>(ix, c, l, r);
      ch[c] = PrefixEnt{(e.code << 2) | c, l, r};
      keep |= (e.l <= e.r && l <= r ? 1u : 0u) << c;
    }
    if (!last) {
      u64 at = block_reserve(static_cast<u32>(__builtin_popcount(keep)), n_out);
#pragma unroll
      for (u32 c = 0; c < 4; ++c)
        if ((keep >> c) & 1) {
          if (at < cap) out[at] = ch[c];
          ++at;
        }
    } else {
#pragma unroll
      for (u32 c = 0; c < 4; ++c) {
        if (!((keep >> c) & 1)) continue;
        // a k-letter interval is one node: its entry takes the first free
        // entry of its home group or of the three groups after it (none: the
        // k-mer stays out and is searched in full). Entries are never freed,
        // so an insert done again finds its own entry before any free one.
        const u64 mix = kmer_hash_mix(ch[c].code, ix.k);
        const u64 home = mix >> rem_bits;
        const u64 node_part = (static_cast<u64>(ch[c].l) + 1) << (64 - node_bits);
        unsigned long long *t64 = reinterpret_cast<unsigned long long *>(table);
        bool placed = false;
        for (u32 d = 0; d < kHashGroups && !placed; ++d) {
          const unsigned long long entry = node_part | kmer_hash_tag(mix, rem_bits, d);
          unsigned long long *group = t64 + 4 * ((home + d) & gmask);
          for (u32 q = 0; q < 4 && !placed; ++q) {
            const unsigned long long prev = atomicCAS(group + q, 0ULL, entry);
            placed = prev == 0 || prev == entry;
          }
        }
      }
    }
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
  std::vector<u64> host(n_blocks * 2, 0);
  parallel_for(n_blocks, [&](u64 b) { host[b * 2 + 1] = get_bits(words, bits, b * kMarkBits, 64); });
  u64 run = 0;
  for (u64 b = 0; b < n_blocks; ++b) {
    host[b * 2] = run;
    run += __builtin_popcountll(host[b * 2 + 1]);
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
  // SBWT_AMD_HYPER_SHIFT (1..25) shrinks the hyperblocks of a wide index so
  // that small test indexes cross many hyperblock boundaries
  hyper_shift = kHyperShift;
  if (const char *hs = std::getenv("SBWT_AMD_HYPER_SHIFT"); wide && hs) {
    const long v = std::strtol(hs, nullptr, 10);
    if (v < 1 || v > static_cast<long>(kHyperShift))
      throw std::runtime_error("SBWT_AMD_HYPER_SHIFT must be in [1, 25]");
    hyper_shift = static_cast<u32>(v);
  }
  if (kp == 0) {
    size_t free_b = 0, total_b = 0;
    SBWT_HIP(hipMemGetInfo(&free_b, &total_b));
    kp = default_presearch_letters(k, n, wide, total_b);
  }
  n_blocks = n / kBlockNodes + 1;  // the line of node n exists: rank(n) is defined
  has_marks = km_words != nullptr;
  // the suffix-group starts are used only when every node outside them has an
  // empty row, as in every SBWT. Rows that break this break the SBWT property
  // the streaming search rests on (a k-mer's successor is one step from a
  // single node), so such an index is searched per k-mer, which is the
  // reference's interval search on whatever rows the file holds
  groups_rejected = sgs_words != nullptr && !suffix_groups_consistent(acgt, sgs_words, n);
  if (groups_rejected) sgs_words = nullptr;
  has_groups = sgs_words != nullptr;
  // line bits; counts are filled by the prefix pass below
  std::vector<uint4> host(n_blocks * kLineSlots);
  const u64 *vecs[6] = {acgt[0], acgt[1], acgt[2], acgt[3], sgs_words, km_words};
  const u64 vec_bits[6] = {n, n, n, n, n, km_words ? km_bits : 0};
  std::vector<u32> last_start(n_blocks);  // last group start in the line + 1 (0: none)
  parallel_for(n_blocks, [&](u64 b) {
    for (u32 s = 0; s < 6; ++s) {
      uint4 e{0, 0, 0, 0};
      if (vecs[s]) {
        u64 lo = get_bits(vecs[s], vec_bits[s], b * kBlockNodes, 64);
        const u64 hi = get_bits(vecs[s], vec_bits[s], b * kBlockNodes + 64, 32);
        if (s == kSlotGroups && b == 0) lo |= 1;  // the root starts the first group
        e.y = static_cast<u32>(lo);
        e.z = static_cast<u32>(lo >> 32);
        e.w = static_cast<u32>(hi);
        if (s == kSlotGroups)
          last_start[b] = hi ? 65 + (31 - __builtin_clz(static_cast<u32>(hi)))
                          : lo ? 1 + (63 - __builtin_clzll(lo)) : 0;
      }
      host[b * kLineSlots + s] = e;
    }
    if (km_words) {  // key-walk codes
      unsigned char code_bytes[32] = {};
      for (u32 o = 0; o < kBlockNodes; ++o) {
        const u64 node = b * kBlockNodes + o;
        u32 code = kWalkNoEdge;
        if (node < n) {
          if (node < km_bits && ((km_words[node >> 6] >> (node & 63)) & 1)) {
            code = kWalkMarked;
          } else {
            for (u32 c = 0; c < 4; ++c)
              if ((acgt[c][node >> 6] >> (node & 63)) & 1) {
                code = c;
                break;
              }
          }
        }
        code_bytes[o / 3] = static_cast<unsigned char>(code_bytes[o / 3] + code * (o % 3 == 0 ? 1 : (o % 3 == 1 ? 6 : 36)));
      }
      std::memcpy(&host[b * kLineSlots + kSlotWalk], code_bytes, 32);
    }
  });
  n_hyper = (n_blocks >> hyper_shift) + 1;
  std::vector<u64> hyper(n_hyper * kLineSlots, 0);
  u64 run[6] = {0, 0, 0, 0, 0, 0};
  u64 group = 0;  // start of the group of the current line's first node
  for (u64 b = 0; b < n_blocks; ++b) {
    for (u32 s = 0; s < 6; ++s) {
      uint4 &e = host[b * kLineSlots + s];
      if (s == kSlotGroups) {
        if (!has_groups) {
          e.x = kFarGroup;
          continue;
        }
        const u64 first = b * kBlockNodes;
        const u64 delta = (e.y & 1) ? 0 : first - group;
        e.x = delta < kFarGroup ? static_cast<u32>(delta) : kFarGroup;
        if (last_start[b]) group = first + last_start[b] - 1;
        continue;
      }
      if ((b & ((1ULL << hyper_shift) - 1)) == 0) hyper[(b >> hyper_shift) * kLineSlots + s] = run[s];
      e.x = static_cast<u32>(wide ? run[s] - hyper[(b >> hyper_shift) * kLineSlots + s] : run[s]);
      run[s] += __builtin_popcount(e.y) + __builtin_popcount(e.z) + __builtin_popcount(e.w);
    }
  }
  c_map[0] = 1;  // src/SbwtBuilder/SbwtBuilder.cpp:111,128-130
  for (u32 c = 0; c < 4; ++c) c_map[c + 1] = c_map[c] + run[c];
  blocks = upload(host.data(), host.size());
  hyper_base = upload(hyper.data(), hyper.size());
  device_bytes = host.size() * sizeof(uint4) + hyper.size() * 8;
  host.clear();
  host.shrink_to_fit();
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
  // key-destination table when the index has key marks and it takes at most
  // an eighth of device memory
  if (has_marks) {
    size_t free_b = 0, total_b = 0;
    SBWT_HIP(hipMemGetInfo(&free_b, &total_b));
    const u64 bytes = n * (wide ? 8 : 4) + 16;
    if (bytes <= total_b / 8 && bytes < free_b / 2) {
      SBWT_HIP(hipMalloc(&key_dest, bytes));
      const IndexView kv = view();
      const u32 grid = static_cast<u32>(std::min<u64>((n + 255) / 256, 1u << 20));
      // SBWT_AMD_KEY_TABLE_STEPS lowers the walk cap so that the query-time
      // walk of long entries can be validated on small indexes
      const char *cap_env = std::getenv("SBWT_AMD_KEY_TABLE_STEPS");
      const u32 cap = cap_env ? static_cast<u32>(std::strtoul(cap_env, nullptr, 10)) : kKeyTableSteps;
      if (wide) k_key_dest<u64><<<grid, 256>>>(kv, static_cast<u64 *>(key_dest), n, cap);
      else k_key_dest<u32><<<grid, 256>>>(kv, static_cast<u32 *>(key_dest), n, cap);
      SBWT_HIP(hipGetLastError());
      SBWT_HIP(hipDeviceSynchronize());
      device_bytes += bytes;
    }
  }
}

// The k-mer table, on request (narrow indexes with k <= 32 and kp < k):
// 2^G groups of four 8-byte entries, G = N (+ 2 at k = 32, so the remainder
// fits the entry) with 2^N > n: at most a quarter of the entries used; built
// when it takes at most half of device memory and leaves an eighth free. The build grows the
// strings of each slice of presearch entries (a range of kp-letter prefixes)
// one letter per level up to k in two level lists; slices are sized so their
// lists fit beside the table, and a slice whose level overflows its list is
// done again as four slices (inserting a k-mer twice leaves one bucket).
// Returns whether the index has one.
bool DeviceIndex::build_kmer_hash() {
  if (kmer_hash) return true;
  if (wide || k > 32 || kp >= k) return false;
  u32 node_bits = 1;  // entry bits of node + 1 <= n
  while (node_bits < 32 && (1ULL << node_bits) <= n) ++node_bits;
  const u32 code_bits = 2 * k;
  const u32 bits = std::min<u32>(code_bits, node_bits + (code_bits > 62 ? code_bits - 62 : 0));
  const u64 table_bytes = (1ULL << bits) * 2 * sizeof(uint4);
  size_t free_b = 0, total_b = 0;
  SBWT_HIP(hipMemGetInfo(&free_b, &total_b));
  const u64 reserve = total_b / 8;
  if (table_bytes > total_b / 2 || table_bytes + reserve >= free_b) return false;
  const u64 count = 1ULL << (2 * kp);
  // list capacity: a whole level when it fits, else what the memory left allows
  const u64 room = (free_b - reserve - table_bytes) / 2 / sizeof(PrefixEnt);
  u64 cap = std::min<u64>(n + 1, room);
  if (cap < std::min<u64>(n + 1, 1u << 16)) return false;  // too little memory left for lists
  u64 first_len = count;  // initial slice: a power of 4 of entries whose share of n fits
  while (first_len > 1 && (n + 1) / (count / first_len) * 3 / 2 > cap) first_len >>= 2;
  // SBWT_AMD_KMER_TABLE_LIST_CAP caps the level lists and starts from one slice,
  // so that slicing and the split of overflowing slices run on small indexes
  if (const char *env = std::getenv("SBWT_AMD_KMER_TABLE_LIST_CAP")) {
    cap = std::max<u64>(1024, std::min<u64>(cap, std::strtoull(env, nullptr, 10)));
    first_len = count;
  }
  const u32 levels = k - kp + 1;
  PrefixEnt *lists[2] = {nullptr, nullptr};
  u64 *counts = nullptr;
  SBWT_HIP(hipMalloc(&lists[0], cap * sizeof(PrefixEnt)));
  SBWT_HIP(hipMalloc(&lists[1], cap * sizeof(PrefixEnt)));
  SBWT_HIP(hipMalloc(&counts, levels * sizeof(u64)));
  SBWT_HIP(hipMalloc(&kmer_hash, table_bytes));
  SBWT_HIP(hipMemset(kmer_hash, 0, table_bytes));
  kmer_hash_bits = bits;
  kmer_hash_node_bits = node_bits;
  const IndexView ix = view();
  std::vector<std::pair<u64, u64>> todo;  // (first entry, entries), done last first
  for (u64 f = count; f > 0;) {
    const u64 len = std::min(first_len, f);
    f -= len;
    todo.emplace_back(f, len);
  }
  std::vector<u64> level_counts(levels);
  try {
    while (!todo.empty()) {
      const auto [first, len] = todo.back();
      todo.pop_back();
      SBWT_HIP(hipMemset(counts, 0, levels * sizeof(u64)));
      k_hash_seed<<<static_cast<u32>(std::min<u64>((len + 256 * kSeedItems - 1) / (256 * kSeedItems), 1u << 20)), 256>>>(
          static_cast<const uint2 *>(presearch), first, len, lists[0], counts, cap);
      SBWT_HIP(hipGetLastError());
      const u32 grid = static_cast<u32>(std::min<u64>((cap + 255) / 256, 1u << 20));
      for (u32 l = kp; l < k; ++l) {
        const u32 from = (l - kp) & 1, to = from ^ 1;
        k_hash_extend<<<grid, 256>>>(ix, lists[from], counts + (l - kp), lists[to],
                                     counts + (l - kp + 1), cap, l + 1 == k ? 1u : 0u, kmer_hash,
                                     bits, node_bits);
        SBWT_HIP(hipGetLastError());
      }
      SBWT_HIP(hipMemcpy(level_counts.data(), counts, levels * sizeof(u64), hipMemcpyDeviceToHost));
      // the last level's entries went to the table; an earlier level that
      // overflowed lost strings, so the slice is done again in quarters
      bool lost = false;
      for (u32 l = 0; l + 1 < levels; ++l) lost |= level_counts[l] > cap;
      if (lost) {
        if (len < 4) throw std::runtime_error("k-mer table: one prefix overflows the level lists");
        for (u32 q = 4; q-- > 0;) todo.emplace_back(first + q * (len / 4), len / 4);
      }
    }
  } catch (...) {
    (void)hipFree(lists[0]);
    (void)hipFree(lists[1]);
    (void)hipFree(counts);
    (void)hipFree(kmer_hash);
    kmer_hash = nullptr;
    kmer_hash_bits = 0;
    throw;
  }
  SBWT_HIP(hipDeviceSynchronize());
  (void)hipFree(lists[0]);
  (void)hipFree(lists[1]);
  (void)hipFree(counts);
  device_bytes += table_bytes;
  return true;
}

void DeviceIndex::release() {
  if (kmer_hash) (void)hipFree(kmer_hash);
  kmer_hash = nullptr;
  kmer_hash_bits = 0;
  if (blocks) (void)hipFree(blocks);
  if (hyper_base) (void)hipFree(hyper_base);
  if (presearch) (void)hipFree(presearch);
  if (key_dest) (void)hipFree(key_dest);
  key_dest = nullptr;
  if (color_table) (void)hipFree(color_table);
  color_table = nullptr;
  color_ranks_in_lines = false;
  if (key_color_masks) (void)hipFree(key_color_masks);
  key_color_masks = nullptr;
  color_table_kind = kColorTableNone;
  blocks = nullptr;
  hyper_base = nullptr;
  presearch = nullptr;
}

IndexView DeviceIndex::view() const {
  IndexView v{};
  v.blocks = blocks;
  v.hyper_base = hyper_base;
  v.hyper_shift = hyper_shift;
  v.presearch = presearch;
  v.key_dest = key_dest;
  v.kmer_hash = kmer_hash;
  v.kmer_hash_bits = kmer_hash_bits;
  v.kmer_hash_node_bits = kmer_hash_node_bits;
  v.has_groups = has_groups;
  v.has_marks = has_marks;
  v.color_table = nullptr;  // searchers that output color-table entries set it
  v.color_table_kind = kColorTableNone;
  for (int i = 0; i < 5; ++i) v.c_map[i] = c_map[i];
  v.num_nodes = n;
  v.k = k;
  v.kp = kp;
  return v;
}

}  // namespace sbwt_amd
