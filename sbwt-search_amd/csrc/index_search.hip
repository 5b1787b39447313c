// Index search kernels for gfx950 and the Searcher that drives them.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "device_fns.hpp"
#include "index_search.hpp"
#include "scan.hpp"

namespace sbwt_amd {

namespace {

constexpr u32 kScanThreads = 256;
constexpr u32 kScanItems = 16;
constexpr u64 kScanChunk = static_cast<u64>(kScanThreads) * kScanItems;
constexpr u64 kMaxGrid = 1u << 20;  // workgroups per launch; kernels grid-stride past it

// ---------------------------------------------------------------------------
// 2-bit packing: src/SeqToBitsConverter/CharToBits.h:21-40 and
// ContinuousSeqToBitsConverter.cpp:90-107. Char j -> word j/32, bits
// [63-2(j%32), 62-2(j%32)]; A/a 0, C/c 1, G/g 2, T/t 3; any other byte packs
// as 0 and sets bit j%32 of invalid word j/32. One thread per output word.
// ---------------------------------------------------------------------------
// Four chars at once (one per byte of bytes4): their four 2-bit codes,
// first char in the top bits of the byte, and their invalid flags (bit q for
// char q). Upper and lower case fold by clearing bit 5; the code of A C G T
// is ((u >> 1) ^ (u >> 2)) & 3; a char is valid iff it equals the letter of
// its code (byte permute from "ACGT"); invalid chars pack as 0.
__device__ __forceinline__ void pack4(u32 bytes4, u32 &codes8, u32 &inv4) {
  const u32 u = bytes4 & 0xDFDFDFDFu;
  u32 code = ((u >> 1) ^ (u >> 2)) & 0x03030303u;
  const u32 letter = __builtin_amdgcn_perm(0u, 0x54474341u, code);  // bytes 'A','C','G','T'
  const u32 diff = letter ^ u;
  const u32 bad = (((diff & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | diff) & 0x80808080u;  // nonzero bytes
  code &= ((~bad >> 7) & 0x01010101u) * 3u;
  // gather the byte fields (disjoint partial products, no carries):
  // c0 << 6 | c1 << 4 | c2 << 2 | c3, and flag q at bit q
  codes8 = (code * 0x40100401u) >> 24;
  inv4 = ((((bad >> 7) & 0x01010101u) * 0x01020408u) >> 24) & 0xFu;
}

__global__ __launch_bounds__(256) void k_pack(const u8 *__restrict__ chars, u64 n, u64 n_words,
                                               bool aligned, u64 *__restrict__ packed,
                                               u32 *__restrict__ invalid) {
  for (u64 w = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; w < n_words;
       w += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 base = w * 32;
    u64 v = 0;
    u32 inv = 0;
    if (aligned && base + 32 <= n) {
      const uint4 a = *reinterpret_cast<const uint4 *>(chars + base);
      const uint4 b = *reinterpret_cast<const uint4 *>(chars + base + 16);
      const u32 g[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      u32 hi = 0, lo = 0;
#pragma unroll
      for (u32 q = 0; q < 8; ++q) {
        u32 c8, i4;
        pack4(g[q], c8, i4);
        if (q < 4) hi |= c8 << (24 - 8 * q);
        else lo |= c8 << (24 - 8 * (q - 4));
        inv |= i4 << (4 * q);
      }
      v = (static_cast<u64>(hi) << 32) | lo;
    } else {
      for (u32 j = 0; j < 32 && base + j < n; ++j) {
        const u32 u = chars[base + j] & 0xDFu;
        const bool ok = (u == 0x41u) | (u == 0x43u) | (u == 0x47u) | (u == 0x54u);
        v |= ok ? (static_cast<u64>(((u >> 1) ^ (u >> 2)) & 3u) << (62 - 2 * j)) : 0ULL;
        inv |= ok ? 0u : (1u << j);
      }
    }
    packed[w] = v;
    invalid[w] = inv;
  }
}

// ---------------------------------------------------------------------------
// Cumulative k-mer counts per read (the positions of
// src/PositionsBuilder/PositionsBuilder.cpp:15-60 made implicit): read r of
// length len has max(0, len - k + 1) k-mers. For the streaming search the
// same pass also accumulates the read's work items, ceil(k-mers / chunk).
// Three-phase scan over blocks of
// kScanChunk reads, each block reading row after row of 256 consecutive
// reads (coalesced) and scanning a row with wave shuffles.
// ---------------------------------------------------------------------------
// (none for a read ending before its start or past the batch's chars)
__device__ __forceinline__ u64 kmers_of_read(const u64 *ends, u64 r, u32 k, u64 n_chars) {
  const u64 st = r ? ends[r - 1] : 0;
  const u64 len = ends[r] >= st && ends[r] <= n_chars ? ends[r] - st : 0;
  return len >= k ? len - k + 1 : 0;
}

struct ReadCounts {
  u64 kmers, items;
};

template <bool ITEMS>
__device__ __forceinline__ ReadCounts read_counts(const u64 *ends, u64 r, u64 n, u32 k, u64 n_chars,
                                                  u32 chunk) {
  ReadCounts c{0, 0};
  if (r >= n) return c;
  c.kmers = kmers_of_read(ends, r, k, n_chars);
  if constexpr (ITEMS) c.items = (c.kmers + chunk - 1) / chunk;
  return c;
}

__device__ __forceinline__ u64 wave_inclusive_sum(u64 v) {
  const u32 lane = threadIdx.x & 63;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u64 y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  return v;
}

// Exclusive prefix and total over the workgroup of one value per thread, in
// thread order; sh holds kScanThreads / 64 entries.
__device__ __forceinline__ u64 block_row_scan(u64 v, u64 *sh, u64 *total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const u64 incl = wave_inclusive_sum(v);
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  u64 before = 0, all = 0;
#pragma unroll
  for (u32 w = 0; w < kScanThreads / 64; ++w) {
    before += w < wave ? sh[w] : 0;
    all += sh[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// sums[blk] = k-mers of the block's reads; with ITEMS, sums[cap + blk] = work items.
template <bool ITEMS>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const u64 *__restrict__ ends, u64 n,
                                                               u32 k, u64 n_chars, u32 chunk,
                                                               u64 *__restrict__ sums, u64 cap) {
  __shared__ u64 sh[2][kScanThreads / 64];
  const u64 base = blockIdx.x * kScanChunk;
  u64 sk = 0, si = 0;
  for (u32 i = 0; i < kScanItems; ++i) {
    const ReadCounts c = read_counts<ITEMS>(ends, base + i * kScanThreads + threadIdx.x, n, k,
                                            n_chars, chunk);
    sk += c.kmers;
    si += c.items;
  }
  u64 tk = 0, ti = 0;
  block_row_scan(sk, sh[0], &tk);
  if constexpr (ITEMS) block_row_scan(si, sh[1], &ti);
  if (threadIdx.x == 0) {
    sums[blockIdx.x] = tk;
    if constexpr (ITEMS) sums[cap + blockIdx.x] = ti;
  }
}

// One workgroup per array (sums + blockIdx.x * cap): in place, exclusive.
__global__ __launch_bounds__(1024) void k_scan_chunks(u64 *__restrict__ sums, u64 n_chunks, u64 cap) {
  __shared__ u64 sh[1024 / 64];
  u64 *a = sums + blockIdx.x * cap;
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u64 carry = 0;
  for (u64 base = 0; base < n_chunks; base += 1024) {
    const u64 i = base + threadIdx.x;
    const u64 v = i < n_chunks ? a[i] : 0;
    const u64 incl = wave_inclusive_sum(v);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    u64 before = 0, all = 0;
    for (u32 w = 0; w < 1024 / 64; ++w) {
      before += w < wave ? sh[w] : 0;
      all += sh[w];
    }
    __syncthreads();
    if (i < n_chunks) a[i] = carry + before + incl - v;
    carry += all;
  }
}

// Inclusive ends: kends[r] = k-mers of reads 0..r; with ITEMS, item_ends[r] =
// work items of reads 0..r.
template <bool ITEMS>
__global__ __launch_bounds__(kScanThreads) void k_scan_write(const u64 *__restrict__ ends, u64 n,
                                                              u32 k, u64 n_chars, u32 chunk,
                                                              const u64 *__restrict__ offs, u64 cap,
                                                              u64 *__restrict__ kends,
                                                              u64 *__restrict__ item_ends) {
  __shared__ u64 sh[2][kScanThreads / 64];
  const u64 base = blockIdx.x * kScanChunk;
  u64 ck = offs[blockIdx.x], ci = ITEMS ? offs[cap + blockIdx.x] : 0;
  for (u32 i = 0; i < kScanItems; ++i) {
    const u64 r = base + i * kScanThreads + threadIdx.x;
    const ReadCounts c = read_counts<ITEMS>(ends, r, n, k, n_chars, chunk);
    u64 tk = 0, ti = 0;
    const u64 bk = block_row_scan(c.kmers, sh[0], &tk);
    if (r < n) kends[r] = ck + bk + c.kmers;
    ck += tk;
    if constexpr (ITEMS) {
      const u64 bi = block_row_scan(c.items, sh[1], &ti);
      if (r < n) item_ends[r] = ci + bi + c.items;
      ci += ti;
    }
  }
}

// One group of the k-mer table (its four entries in g0, g1) read for the
// k-mer whose entry there would end in `tag` (its low 64 - node_bits bits):
// 1 its entry (node set), 0 an empty entry first (the k-mer is absent), -1
// four entries of other k-mers (read the next group).
__device__ __forceinline__ int probe_group(const uint4 g0, const uint4 g1, u64 tag, u32 node_bits,
                                           u32 &node) {
  const u64 e[4] = {static_cast<u64>(g0.x) | (static_cast<u64>(g0.y) << 32),
                    static_cast<u64>(g0.z) | (static_cast<u64>(g0.w) << 32),
                    static_cast<u64>(g1.x) | (static_cast<u64>(g1.y) << 32),
                    static_cast<u64>(g1.z) | (static_cast<u64>(g1.w) << 32)};
  const u64 low = ~0ULL >> node_bits;
#pragma unroll
  for (u32 q = 0; q < 4; ++q) {
    if (e[q] == 0) return 0;
    if ((e[q] & low) == tag) {
      node = static_cast<u32>(e[q] >> (64 - node_bits)) - 1;
      return 1;
    }
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Search of one k-mer whose 2k bits sit at the top of `window`:
// src/IndexSearcher/IndexSearcher.cuh:36-71. The presearch table gives the
// interval after kp letters; each remaining letter is one SBWT step. Once the
// interval is empty it stays empty (l = r + 1 is preserved by the step), so
// the loop stops there: the result is the same not-found sentinel.
// ---------------------------------------------------------------------------
// `steps` and `walked` count the SBWT steps and key-walk steps taken.
template <typename Idx, bool KEY>
__device__ __forceinline__ u64 search_window(const IndexView &ix, u64 window, u32 &steps,
                                             u32 &walked) {
  if (ix.kmer_hash) {  // the k-mer table answers in a group or two
    const u64 mix = kmer_hash_mix(window >> (64 - 2 * ix.k), ix.k);
    const u32 rem_bits = 2 * ix.k - ix.kmer_hash_bits;
    const u64 gmask = (1ULL << ix.kmer_hash_bits) - 1;
    u64 g = mix >> rem_bits;
    // `steps` counts the further lines read, the first standing in for the
    // presearch line the caller counts
    for (u32 d = 0; d < kHashGroups; ++d, g = (g + 1) & gmask) {
      const uint4 *group = ix.kmer_hash + 2 * g;
      if (d && (g & 3) == 0) ++steps;  // the group opens a new line
      u32 node = 0;
      const int pr = probe_group(group[0], group[1], kmer_hash_tag(mix, rem_bits, d),
                                 ix.kmer_hash_node_bits, node);
      if (pr == 1) {
        if constexpr (KEY) return key_result<Idx>(ix, static_cast<Idx>(node), walked);
        return static_cast<u64>(node);
      }
      if (pr == 0) return kNotFound;
    }
    ++steps;  // the lines read were all extra: the presearch line follows
  }
  const u64 pre = window >> (64 - 2 * ix.kp);
  Idx l, r;
  if constexpr (sizeof(Idx) == 4) {
    const uint2 p = static_cast<const uint2 *>(ix.presearch)[pre];
    l = p.x;
    r = p.y;
  } else {
    const ulonglong2 p = static_cast<const ulonglong2 *>(ix.presearch)[pre];
    l = p.x;
    r = p.y;
  }
  for (u32 j = ix.kp; j < ix.k && l <= r; ++j) {
    const u32 c = static_cast<u32>(window >> (62 - 2 * j)) & 3u;
    sbwt_step<Idx>(ix, c, l, r);
    ++steps;
  }
  if (l > r) return kNotFound;
  if constexpr (KEY) {
    return key_result<Idx>(ix, l, walked);
  } else {
    return static_cast<u64>(l);
  }
}

__device__ __forceinline__ u64 window_at(const u64 *__restrict__ packed, u64 p) {
  const u64 w = p >> 5;
  const u32 sh = static_cast<u32>(p & 31) * 2;
  const u64 a = packed[w];
  return sh ? ((a << sh) | (packed[w + 1] >> (64 - sh))) : a;
}

// Reference interface: explicit positions (IndexSearcher.cuh:22-72).
template <typename Idx, bool KEY>
__global__ __launch_bounds__(kSearchThreads) void k_search_positions(
    IndexView ix, const u64 *__restrict__ packed, const u64 *__restrict__ positions, u64 n,
    u64 *__restrict__ out) {
  u32 steps = 0, walked = 0;
  for (u64 i = static_cast<u64>(blockIdx.x) * kSearchThreads + threadIdx.x; i < n;
       i += static_cast<u64>(gridDim.x) * kSearchThreads) {
    out[i] = search_window<Idx, KEY>(ix, window_at(packed, positions[i]), steps, walked);
  }
}

// Fused reads form: workgroup w handles k-mers [256 w, 256 w + 256). Threads 0
// and 255 locate the reads of the first and last k-mer by binary search over
// the cumulative k-mer counts; every lane then searches that short range. A
// k-mer with an invalid char is kInvalid without a search
// (ContinuousIndexResultsPrinter.hpp:199-223, 289-308). With STATS the kernel
// also adds {k-mers searched, SBWT steps, key-walk steps, invalid k-mers} into
// stats[0..3] (a census run for the algorithmic byte count; same results).
template <typename Idx, bool KEY, bool STATS>
__global__ __launch_bounds__(kSearchThreads) void k_search_reads(
    IndexView ix, const u64 *__restrict__ packed, const u32 *__restrict__ invalid,
    const u64 *__restrict__ char_ends, const u64 *__restrict__ kmer_ends, u64 n_reads,
    u64 n_chars, u64 *__restrict__ out, u64 *__restrict__ stats) {
  __shared__ u64 sh_r[2];
  __shared__ u64 sh_stats[4];
  if (STATS && threadIdx.x < 4) sh_stats[threadIdx.x] = 0;
  u32 steps = 0, walked = 0, searched = 0, invalids = 0;
  // k-mers never outnumber chars; the bound keeps malformed read ends in range
  const u64 n_kmers = min(kmer_ends[n_reads - 1], n_chars);
  const u64 n_tiles = (n_kmers + kSearchThreads - 1) / kSearchThreads;
  const u32 t = threadIdx.x;
  const u64 kmask = (ix.k >= 64) ? ~0ULL : ((1ULL << ix.k) - 1);
  for (u64 tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const u64 g0 = tile * kSearchThreads;
    if (t == 0 || t == kSearchThreads - 1) {
      const u64 g = t == 0 ? g0 : min(g0 + kSearchThreads - 1, n_kmers - 1);
      u64 lo = 0, hi = n_reads;
      while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (kmer_ends[mid] > g) hi = mid;
        else lo = mid + 1;
      }
      sh_r[t == 0 ? 0 : 1] = lo;
    }
    __syncthreads();
    const u64 g = g0 + t;
    u64 lo = sh_r[0], hi = sh_r[1];
    __syncthreads();
    if (g < n_kmers) {
      while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (kmer_ends[mid] > g) hi = mid;
        else lo = mid + 1;
      }
      const u64 r = lo;
      const u64 kstart = r ? kmer_ends[r - 1] : 0;
      const u64 cstart = r ? char_ends[r - 1] : 0;
      const u64 p = cstart + (g - kstart);
      if (p + ix.k > n_chars) {
        out[g] = kInvalid;
        continue;
      }
      const u64 iw = p >> 5;
      const u64 ib = ((static_cast<u64>(invalid[iw + 1]) << 32) | invalid[iw]) >> (p & 31);
      if (ib & kmask) {
        out[g] = kInvalid;
        ++invalids;
      } else {
        out[g] = search_window<Idx, KEY>(ix, window_at(packed, p), steps, walked);
        ++searched;
      }
    }
  }
  if constexpr (STATS) {
    __syncthreads();
    atomicAdd(reinterpret_cast<unsigned long long *>(&sh_stats[0]), static_cast<u64>(searched));
    atomicAdd(reinterpret_cast<unsigned long long *>(&sh_stats[1]), static_cast<u64>(steps));
    atomicAdd(reinterpret_cast<unsigned long long *>(&sh_stats[2]), static_cast<u64>(walked));
    atomicAdd(reinterpret_cast<unsigned long long *>(&sh_stats[3]), static_cast<u64>(invalids));
    __syncthreads();
    if (threadIdx.x < 4)
      atomicAdd(reinterpret_cast<unsigned long long *>(stats + threadIdx.x), sh_stats[threadIdx.x]);
  }
}

// ---------------------------------------------------------------------------
// Streaming search with persistent lanes. Work items are chunks of up to
// kChunkKmers consecutive k-mers of one read. A lane walks its chunk k-mer by
// k-mer; when k-mer x was found at node v, the next k-mer y = x[1..k-1] c is
// one SBWT step from the suffix group of v: if v carries edge c, y is node
// C[c] + rank_c(v) (one rank-block line); if v starts its suffix group and
// lacks c, y is absent; otherwise the group start gs (the last set bit of the
// suffix-group starts at or before v, which carries the group's edges)
// decides the same way. A k-mer that follows an absent, invalid or
// chunk-initial one is searched from its presearch interval as in
// search_window. Results equal the per-k-mer search: each k-mer's node is
// unique. The chunk length (k-mers per work item) and the number of
// persistent workgroups per CU are launch parameters (set_launch, autotune).
// ---------------------------------------------------------------------------

__device__ __forceinline__ u32 lanes_below(u64 mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<u32>(mask), 0u));
}

// Per read: its chunks {first k-mer (output slot), char position | count << 48}
// (a read ending past n_chars has none, so malformed read ends never send the
// search outside the batch: the scan counted it no work items).
__global__ __launch_bounds__(256) void k_chunk_fill(const u64 *__restrict__ kmer_ends,
                                                     const u64 *__restrict__ char_ends, u64 n_reads,
                                                     const u64 *__restrict__ chunk_ends, u32 chunk,
                                                     u64 n_out, ulonglong2 *__restrict__ chunks) {
  for (u64 r = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; r < n_reads;
       r += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 ks = r ? kmer_ends[r - 1] : 0, nk = kmer_ends[r] - ks;
    const u64 cs = r ? char_ends[r - 1] : 0;
    u64 at = r ? chunk_ends[r - 1] : 0;
    for (u64 off = 0; at < chunk_ends[r]; off += chunk, ++at) {
      const u64 cnt = min(static_cast<u64>(chunk), nk - off);
      ulonglong2 ch;
      // output slots stay below n_out even for read ends that go back and
      // forth (which break the read-ends contract; their results are void)
      ch.x = min(ks + off, n_out - cnt);
      ch.y = (cs + off) | (cnt << 48);
      chunks[at] = ch;
    }
  }
}

__global__ void k_reset_counter(u64 *c) { *c = 0; }

// Every loop iteration, each lane issues exactly two independent 16-byte
// loads whose addresses depend on its state, so a wave pays one memory round
// trip per iteration however its lanes diverge:
//   kChunk       the chunk's descriptor (output slot, char position, count)
//   kSeq0/kSeq1  64 packed chars + their 64 invalid bits (a lane keeps 128
//                chars in registers; k-mer windows are shifted out of them)
//   kStream      letter slot c and suffix-group slot of v's node line (one line)
//   kGsEntry     letter slot c of the line holding v's group start, when the
//                group began in an earlier line
//   kPresearch   presearch interval of the k-mer's first kp letters
//   kStep        letter slot c of the lines of l and of r + 1
// Output: a lane's results are consecutive slots of out. They are staged in
// LDS (one column per lane, 8 results) and written as whole 64-byte groups
// with four 16-byte stores; partial groups at chunk ends go out one by one.
// Scattered 8-byte stores from every lane would keep more partly written
// lines open than the L2 holds, and each eviction would write a fragment.
// Work distribution: each wave owns a range of chunks and hands them to its
// lanes as they finish (ballot + popcount, no atomics). The first half of the
// chunks is split evenly over the waves up front; the rest is taken in ranges
// of 64 from a device-wide counter, one atomic per range issued beside the
// loads. Device-scope atomics on one address serialize at the memory side, so
// a counter touched once per lane refill stalls every wave; per 64 chunks it
// only balances the tail.
// Steps per loop iteration at most: the first through the full state
// machine, the others stream steps settled from the lane's node line alone (an
// edge of the node, or of its suffix-group start in the same line, or a miss)
// and k-mer table probes. Each pass adds a dependent round trip to the
// iteration; on C2, 2 / 3 / 4 / 5 / 6 passes: search kernel 23.2 / 22.0 /
// 21.6 / 21.5 / 21.4 ms (C3's index stage 32.4 / 32.6 / 32.3 ms for 2-4).
constexpr u32 kStreamPasses = 4;
// Without a k-mer table, fresh searches step through the SBWT one full
// iteration per letter, and every pass lengthens those iterations: on C2
// without the table, 1 / 2 / 3 / 4 passes: search kernel 30.0 / 29.2 / 30.8 /
// 32.8 ms.
constexpr u32 kStreamPassesNoTable = 2;

// v's suffix-group start when its line holds it: the last group-start bit at
// or before offset o in the line's group slot dg (pos = its offset); false
// when the group began in an earlier line.
__device__ __forceinline__ bool group_start_in_line(const uint4 dg, u32 o, u32 &pos) {
  const u32 q = o + 1;
  const u64 glo = (static_cast<u64>(dg.y) | (static_cast<u64>(dg.z) << 32)) &
                  (q >= 64 ? ~0ULL : ((1ULL << q) - 1));
  const u32 ghi = q > 64 ? (dg.w & ((q - 64 >= 32) ? ~0u : ((1u << (q - 64)) - 1))) : 0u;
  if ((ghi | glo) == 0) return false;
  pos = ghi ? 95 - static_cast<u32>(__builtin_clz(ghi)) : 63 - static_cast<u32>(__builtin_clzll(glo));
  return true;
}

// The 32-bit word of line b's walk slots that holds the key-walk code of the
// node at offset o: its three-node group's byte, four groups per word.
__device__ __forceinline__ const u32 *walk_word(const IndexView &ix, u64 b, u32 o) {
  const u32 o3 = o >= kWalkNodesPerSlot ? o - kWalkNodesPerSlot : o;
  return reinterpret_cast<const u32 *>(line_slot(ix, b, kSlotWalk + (o >= kWalkNodesPerSlot ? 1u : 0u))) +
         (o3 / 3 >> 2);
}

// What v's key-walk code (in `word`, read from walk_word of v's line and
// offset o) says about the k-mer found on v's edge c: v marked, v with no edge
// (stuck), v's first edge is c (the two k-mers share a key), or unknown.
__device__ __forceinline__ u32 walk_pred(u32 word, u32 o, bool has_c, u32 c) {
  // v's code: the byte of its three-node group, then its base-6 digit
  const u32 o3 = o >= kWalkNodesPerSlot ? o - kWalkNodesPerSlot : o;
  const u32 q3 = o3 / 3;
  const u32 byte = (word >> (8 * (q3 & 3))) & 255u;
  const u32 dig = o3 - 3 * q3;
  const u32 wcode = (dig == 0 ? byte : (dig == 1 ? byte / 6 : byte / 36)) % 6;
  return wcode == kWalkMarked ? kPredMarked
         : (wcode == kWalkNoEdge ? kPredStuck : (has_c && wcode == c ? kPredSameKey : kPredLookup));
}

template <typename Idx, bool KEY, bool STATS>
__global__ __launch_bounds__(kSearchThreads) void k_search_stream(
    IndexView ix, const u64 *__restrict__ packed, const u32 *__restrict__ invalid,
    const ulonglong2 *__restrict__ chunks, const u64 *__restrict__ chunk_ends, u64 n_reads,
    u64 *__restrict__ counter, u64 *__restrict__ out, u64 *__restrict__ stats, u32 absence) {
  enum : u32 { kChunk, kSeq0, kSeq1, kStream, kGsEntry, kPresearch, kTail, kStep, kHash, kDone };
  enum : u32 { kHintNone, kHintMiss, kHintPre };  // what the last miss left known
  const u64 n_chunks = chunk_ends[n_reads - 1];
  const u32 lane = threadIdx.x & 63;
  const u32 k = ix.k, kp = ix.kp;
  const u64 inv_mask = (1ULL << k) - 1;
  const uint4 *packed4 = reinterpret_cast<const uint4 *>(packed);   // 64 chars each
  const uint4 *invalid4 = reinterpret_cast<const uint4 *>(invalid); // 128 chars each
  const uint4 *pre4 = reinterpret_cast<const uint4 *>(ix.presearch);
  const uint4 *chunks4 = reinterpret_cast<const uint4 *>(chunks);
  const uint4 *dummy = ix.blocks;
  // the wave's unassigned chunks [wnext, wend); the first half of all chunks
  // is split evenly, the rest comes from the counter (offset by that half)
  const u64 n_waves = static_cast<u64>(gridDim.x) * (kSearchThreads / 64);
  const u64 wave = static_cast<u64>(blockIdx.x) * (kSearchThreads / 64) + threadIdx.x / 64;
  const u64 share = n_chunks / (2 * n_waves);
  const u64 dyn0 = share * n_waves;
  u64 wnext = wave * share, wend = wnext + share;
  bool dry = false;  // the counter is past the last chunk
  u32 mode = kChunk;
  bool want = true;  // the lane waits for a chunk
  u64 g = 0, p = 0, buf = 0;           // output slot, char position, buffer start (64-aligned)
  u64 w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // packed chars [buf, buf + 128)
  u64 i0 = 0, i1 = 0;                  // their invalid bits
  u64 window = 0, pre_idx = 0;       // current window; presearch entry to read
  // windows starting before dead_end are not found; in kTail, the checked
  // kp-mer's start (every window before it is past dead_end then)
  u64 dead_end = 0;
  u64 miss_at = 0;                    // char that ended the last miss
  u32 hint = kHintNone;
  u32 left = 0, j = 0, c = 0;
  // KEY: what the next found result says about v (set by kStream): kPred* in
  // bits 0-1, 1 + v's rank among the marked nodes above them (or 0)
  u32 pred = kPredLookup;
  Idx l = 0, r = 0, v = 0;  // interval; found node; l: also v's group start in kGsEntry
  bool prev = false;
  __shared__ u64 obuf[8][kSearchThreads];  // staged results, obuf[slot][thread]
  const u32 tid = threadIdx.x;
  const u64 phase = (reinterpret_cast<uintptr_t>(out) >> 3) & 7;  // out's offset in a 64-byte group
  u64 gbeg = 0;  // first staged output slot
  u32 st_full = 0, st_steps = 0, st_walk = 0, st_inv = 0, st_stream = 0, st_extra = 0, st_fb = 0;
  u32 st_group = 0, st_tail = 0, st_tail_hit = 0, st_skip = 0, st_hash = 0;
  const u64 hash_mask = ix.kmer_hash ? (1ULL << ix.kmer_hash_bits) - 1 : 0;
  const u32 passes = ix.kmer_hash ? kStreamPasses : kStreamPassesNoTable;  // steps per iteration
  const u32 rem_bits = 2 * k - ix.kmer_hash_bits;
  // in kHash, pre_idx = the k-mer's mix and j = the groups read before this
  // one: the group to read, and the low bits of the k-mer's entry there
  auto hash_group = [&]() { return ((pre_idx >> rem_bits) + j) & hash_mask; };
  auto hash_tag = [&]() { return kmer_hash_tag(pre_idx, rem_bits, j); };
  // a fresh search of the current window: the k-mer table when there is one,
  // else the presearch entry and steps
  auto start_search = [&]() {
    if (ix.kmer_hash) {
      pre_idx = kmer_hash_mix(window >> (64 - 2 * k), k);
      j = 0;
      mode = kHash;
    } else {
      pre_idx = window >> (64 - 2 * kp);
      mode = kPresearch;
    }
  };
  for (;;) {
    // hand the wave's chunks to lanes that wait for one
    u64 ci = 0;  // the lane's chunk, read in this iteration's kChunk
    const u64 need = __ballot(want);
    if (need != 0) {
      const u64 avail = wend - wnext;
      if (want && lanes_below(need) < avail) {
        ci = wnext + lanes_below(need);
        want = false;
      } else if (want && dry) {
        want = false;
        mode = kDone;
      }
      wnext += min(static_cast<u64>(__popcll(need)), avail);
    }
    if (__ballot(mode != kDone) == 0) break;
    // an empty range is refilled by one atomic, answered with this iteration's loads
    const bool refill = !dry && wnext == wend;
    u64 tbase = 0;
    if (refill && lane == 0)
      tbase = atomicAdd(reinterpret_cast<unsigned long long *>(counter), 64ULL);
    // ---- the two loads of this iteration
    // nodes of the entry loads (0 in other states: indexes stay in range)
    const Idx n1 = mode == kStream ? v : (mode == kGsEntry || mode == kStep ? l : Idx(0));
    const Idx n2 = mode == kStep ? static_cast<Idx>(r + 1) : Idx(0);
    const Idx b1 = block_of(n1), b2 = block_of(n2);
    const u32 o1 = static_cast<u32>(n1 - b1 * kBlockNodes), o2 = static_cast<u32>(n2 - b2 * kBlockNodes);
    const uint4 *e1 = line_slot(ix, b1, c);
    const uint4 *a1 = dummy, *a2 = dummy;
    switch (mode) {
      case kChunk:
        if (!want) a1 = a2 = chunks4 + ci;
        break;
      case kSeq0:
        a1 = packed4 + (buf >> 6);
        a2 = invalid4 + (buf >> 7);
        break;
      case kSeq1:
        a1 = packed4 + (buf >> 6) + 1;
        a2 = invalid4 + ((buf + 64) >> 7);
        break;
      case kStream:
        a1 = e1;
        a2 = ix.has_groups ? line_slot(ix, b1, kSlotGroups) : e1;
        break;
      case kGsEntry:
        a1 = a2 = e1;
        break;
      case kPresearch:
      case kTail:
        a1 = a2 = pre4 + (sizeof(Idx) == 4 ? (pre_idx >> 1) : pre_idx);
        break;
      case kStep:
        a1 = e1;
        a2 = line_slot(ix, b2, c);
        break;
      case kHash:  // a group of four entries
        a1 = ix.kmer_hash + 2 * hash_group();
        a2 = a1 + 1;
        break;
      default:
        break;
    }
    uint4 d1 = *a1;
    uint4 d2 = *a2;
    // keep the loaded words in registers: re-loading them later in a branch
    // would add a dependent round trip
    asm volatile("" : "+v"(d1.x), "+v"(d1.y), "+v"(d1.z), "+v"(d1.w), "+v"(d2.x), "+v"(d2.y),
                 "+v"(d2.z), "+v"(d2.w));
    // KEY: the word of v's key-walk code, from the line the stream step reads anyway
    // and, with line ranks, v's marks slot: v's rank among the marked nodes
    u32 d3 = 0, mrank = 0;
    if constexpr (KEY) {
      d3 = *(mode == kStream ? walk_word(ix, b1, o1) : reinterpret_cast<const u32 *>(dummy));
      uint4 dm = *(mode == kStream && ix.line_ranks ? line_slot(ix, b1, kSlotMarks) : dummy);
      asm volatile("" : "+v"(d3), "+v"(dm.x), "+v"(dm.y), "+v"(dm.z), "+v"(dm.w));
      mrank = dm.x + entry_popc(dm, o1);
    }
    if (refill) {
      tbase = dyn0 + __shfl(tbase, 0, 64);
      if (tbase >= n_chunks) {
        dry = true;
      } else {
        wnext = tbase;
        wend = min(tbase + 64, n_chunks);
      }
    }
    // ---- rank arithmetic shared by the entry states (kStream, kGsEntry, kStep)
    const Idx cm = cmap_of<Idx>(ix, c);
    const Idx x1 = cm + entry_base(ix, d1, b1, c) + entry_popc(d1, o1);
    const bool bit1 = entry_bit(d1, o1);
    // ---- advance the lane's state
    bool found = false, notfound = false, decide = false, resolve = false;
    Idx found_at = x1;
    switch (mode) {
      case kChunk:
        if (!want) {
          g = static_cast<u64>(d1.x) | (static_cast<u64>(d1.y) << 32);
          gbeg = g;
          const u64 y = static_cast<u64>(d1.z) | (static_cast<u64>(d1.w) << 32);
          p = y & ((1ULL << 48) - 1);
          left = static_cast<u32>(y >> 48);
          buf = p & ~63ULL;
          prev = false;
          hint = kHintNone;
          dead_end = 0;
          mode = kSeq0;
        }
        break;
      case kSeq0:
        w0 = static_cast<u64>(d1.x) | (static_cast<u64>(d1.y) << 32);
        w1 = static_cast<u64>(d1.z) | (static_cast<u64>(d1.w) << 32);
        i0 = (buf & 64) ? (static_cast<u64>(d2.z) | (static_cast<u64>(d2.w) << 32))
                        : (static_cast<u64>(d2.x) | (static_cast<u64>(d2.y) << 32));
        mode = kSeq1;
        break;
      case kSeq1:
        w2 = static_cast<u64>(d1.x) | (static_cast<u64>(d1.y) << 32);
        w3 = static_cast<u64>(d1.z) | (static_cast<u64>(d1.w) << 32);
        i1 = ((buf + 64) & 64) ? (static_cast<u64>(d2.z) | (static_cast<u64>(d2.w) << 32))
                               : (static_cast<u64>(d2.x) | (static_cast<u64>(d2.y) << 32));
        decide = true;
        break;
      case kStream:
        if (STATS) ++st_stream;
        if constexpr (KEY) {
          // what v's line says about v (kPred*), and, when v is marked and the
          // color table ranks keys like the lines' marks, 1 + v's rank above it
          pred = walk_pred(d3, o1, bit1, c);
          if (pred == kPredMarked && ix.line_ranks && mrank <= kRankMax) pred |= (mrank + 1) << 2;
        }
        if (bit1) {
          found = true;
        } else if (!ix.has_groups) {
          if (STATS) ++st_fb;
          start_search();  // the group start is unknown: search afresh
        } else {
          // v's group start: the last group-start bit at or before v in this
          // line, else delta nodes before the line
          u32 pos = 0;
          if (group_start_in_line(d2, o1, pos)) {
            if (STATS && pos != o1) ++st_group;
            // the group's edges sit at its start (v itself when pos == o1)
            if (pos != o1 && entry_bit(d1, pos)) {
              found_at = cm + entry_base(ix, d1, b1, c) + entry_popc(d1, pos);
              found = true;
            } else {
              notfound = true;
            }
          } else if (d2.x == kFarGroup) {
            if (STATS) ++st_fb;
            start_search();
          } else {
            l = static_cast<Idx>(static_cast<u64>(b1) * kBlockNodes - d2.x);
            if (STATS) ++st_group;
            mode = kGsEntry;
          }
        }
        break;
      case kGsEntry:
        if (STATS) ++st_extra;
        found = bit1;
        notfound = !bit1;
        break;
      case kPresearch:
      case kTail:
        if constexpr (sizeof(Idx) == 4) {
          l = (pre_idx & 1) ? d1.z : d1.x;
          r = (pre_idx & 1) ? d1.w : d1.y;
        } else {
          l = static_cast<Idx>(static_cast<u64>(d1.x) | (static_cast<u64>(d1.y) << 32));
          r = static_cast<Idx>(static_cast<u64>(d1.z) | (static_cast<u64>(d1.w) << 32));
        }
        if (mode == kTail) {
          // absence check: the kp-mer at dead_end lies in every window [p, dead_end]
          if (STATS) ++st_tail;
          if (l > r) {
            if (STATS) ++st_tail_hit;
            dead_end += 1;
            notfound = true;
          } else {
            dead_end = 0;
            start_search();
          }
          break;
        }
        j = kp;
        if (STATS) ++st_full;
        mode = kStep;
        resolve = true;
        break;
      case kStep:
        l = x1;
        r = cm + entry_base(ix, d2, b2, c) + entry_popc(d2, o2) - 1;
        ++j;
        if (STATS) ++st_steps;
        resolve = true;
        break;
      case kHash: {
        u32 node = 0;
        const int pr = probe_group(d1, d2, hash_tag(), ix.kmer_hash_node_bits, node);
        const bool settled = pr >= 0;
        if (pr == 1) {
          found_at = static_cast<Idx>(node);
          found = true;
        } else if (pr == 0) {
          notfound = true;
        } else {
          if (++j < kHashGroups) {
            if (STATS && (hash_group() & 3) == 0) ++st_hash;  // the next group opens a new line
          } else {  // a long probe run: the full search answers
            if (STATS) ++st_hash;  // (its presearch line is the search's own)
            pre_idx = window >> (64 - 2 * kp);
            mode = kPresearch;
          }
        }
        if (settled) {
          if (STATS) ++st_full;  // the first line stands in for the presearch line of a full search
          if (notfound) {  // where the k-mer fails is unknown: check its last kp-mer next
            hint = kHintMiss;
            miss_at = p + k - 1;
          }
        }
        break;
      }
      default:
        break;
    }
    if (resolve) {  // a search just started or stepped
      if (l > r) {
        notfound = true;
        // the miss ended at char p + j - 1, or within the presearch prefix
        hint = j == kp ? kHintPre : kHintMiss;
        miss_at = p + j - 1;
      } else if (j == k) {
        found_at = l;
        found = true;
      } else {
        c = static_cast<u32>(window >> (62 - 2 * j)) & 3u;
      }
    }
    u64 res = kNotFound;
    bool emitted = notfound;
    if (notfound && (mode == kStream || mode == kGsEntry)) {
      // the k-mer's last char made the miss: its predecessor was found
      hint = kHintMiss;
      miss_at = p + k - 1;
    }
    if (notfound) {
      prev = false;
      pred = kPredLookup;
    }
    if (found) {
      v = found_at;
      prev = true;
      hint = kHintNone;
      res = static_cast<u64>(v);
      if constexpr (KEY) {
        // the node and what it says about the k-mer before it; k_key_runs
        // replaces it with the key node
        res |= (static_cast<u64>(pred & 3u) << kPredShift) | (static_cast<u64>(pred >> 2) << kRankShift);
        pred = kPredLookup;
      }
      emitted = true;
    }
    // ---- emit, then set up the next k-mer (invalid k-mers resolve here)
    // stage one result in LDS; a complete 64-byte group (or the chunk's end)
    // goes out to HBM
    auto stage = [&](u64 res_) {
      const u32 slot = static_cast<u32>((g + phase) & 7);
      obuf[slot][tid] = res_;
      if (slot == 7 || left == 1) {
        if (slot == 7 && g - gbeg == 7) {
          uint4 *o4 = reinterpret_cast<uint4 *>(out + gbeg);  // 64-byte aligned
#pragma unroll
          for (u32 q = 0; q < 4; ++q) {
            const u64 a = obuf[2 * q][tid], b = obuf[2 * q + 1][tid];
            o4[q] = make_uint4(static_cast<u32>(a), static_cast<u32>(a >> 32), static_cast<u32>(b),
                               static_cast<u32>(b >> 32));
          }
        } else {
          for (u64 i = gbeg; i <= g; ++i) out[i] = obuf[(i + phase) & 7][tid];
        }
        gbeg = g + 1;
      }
      ++g;
      ++p;
      --left;
    };
    // emit res_ (when emitted_), then set up the lane's next k-mer: its window,
    // and the state that finds it
    auto advance = [&](u64 res_, bool emitted_, bool decide_) {
    while (emitted_ || decide_) {
      if (emitted_) {
        stage(res_);
        emitted_ = false;
        if (left == 0) {  // chunk done: wait for the next one
          want = true;
          mode = kChunk;
          break;
        }
      }
      decide_ = false;
      // windows settled by an earlier absence check: a tight loop, since a
      // run (up to k - kp + 1 windows) would otherwise hold the whole wave in
      // this loop's full body once per window
      while (p < dead_end && p - buf < 64 && left > 0) {
        const u32 q = static_cast<u32>(p - buf);
        const bool bad = ((q ? ((i0 >> q) | (i1 << (64 - q))) : i0) & inv_mask) != 0;
        if (STATS) {
          if (bad) ++st_inv;
          else ++st_skip;
        }
        stage(bad ? kInvalid : kNotFound);
      }
      if (left == 0) {
        want = true;
        mode = kChunk;
        break;
      }
      if (p - buf >= 64) {  // slide the 128-char buffer by 64 chars
        buf += 64;
        w0 = w2;
        w1 = w3;
        i0 = i1;
        mode = kSeq1;
        break;
      }
      const u32 q = static_cast<u32>(p - buf);
      const u64 wa = (q & 32) ? w1 : w0, wb = (q & 32) ? w2 : w1;
      const u32 sh = 2 * (q & 31);
      window = sh ? ((wa << sh) | (wb >> (64 - sh))) : wa;
      const u64 iv = q ? ((i0 >> q) | (i1 << (64 - q))) : i0;
      if (iv & inv_mask) {
        res_ = kInvalid;
        prev = false;
        hint = kHintNone;
        emitted_ = true;
        if (STATS) ++st_inv;
        continue;
      }
      if (prev) {
        c = static_cast<u32>(window >> (64 - 2 * k)) & 3u;
        mode = kStream;
      } else {
        // Absence check before a full search: the kp-mer at t lies in every
        // window from p to t, so if its presearch interval is empty all of
        // them are not found. After a miss that ended at char miss_at the
        // check takes the last kp-mer of this window that still holds that
        // char; with nothing known, the window's last kp-mer. After a miss
        // in the presearch prefix the next window's prefix likely holds the
        // same bad char, so it is searched directly.
        u64 t = p;
        // with the k-mer table a search costs about what a check does: only a
        // known miss is worth checking around
        if (absence && hint == kHintNone && !ix.kmer_hash) t = p + k - kp;
        if (absence && hint == kHintMiss && miss_at >= p) t = min(miss_at, p + k - kp);
        if (t > p) {
          dead_end = t;
          const u32 x = static_cast<u32>(t - buf);  // < 96: the kp-mer ends inside this window
          const u64 ta = x < 32 ? w0 : (x < 64 ? w1 : w2), tb = x < 32 ? w1 : (x < 64 ? w2 : w3);
          const u32 tsh = 2 * (x & 31);
          pre_idx = (tsh ? ((ta << tsh) | (tb >> (64 - tsh))) : ta) >> (64 - 2 * kp);
          mode = kTail;
        } else {
          start_search();
        }
      }
    }
    };
    // Then further steps in the same iteration (`passes` in all) for the
    // two commonest states: a stream step that v's own line settles (the next
    // k-mer on v's edge c, on the edge c of v's suffix-group start in the same
    // line, or a miss), and a k-mer table probe. Every other state's code then
    // runs once per several k-mers of most lanes, not per one; any other
    // outcome (a group that began in an earlier line, a long probe run) waits
    // for the next iteration's full step. Absence checks stay in the full step:
    // settled there, their runs of skipped windows (up to k - kp + 1) would run
    // once per pass (C2: 22.0 ms without them in the passes, 27.1 ms with).
    // One copy of advance() serves every pass (two copies cost registers, then
    // spills).
#pragma clang loop unroll(disable)
    for (u32 pass = 0; pass < passes; ++pass) {
      advance(res, emitted, decide);
      emitted = false;
      decide = false;
      bool moved = false;
      if (pass + 1 < passes && mode == kStream) {
        const Idx bs = block_of(v);
        const u32 os = static_cast<u32>(v - bs * kBlockNodes);
        const uint4 ds = *line_slot(ix, bs, c);
        // the group slot of the same line (v's suffix-group start), and, KEY,
        // v's key-walk code, as in the full step
        const uint4 dg = *line_slot(ix, bs, ix.has_groups ? kSlotGroups : c);
        u32 dw = 0;
        if constexpr (KEY) dw = *walk_word(ix, bs, os);
        const bool has_c = entry_bit(ds, os);
        u32 at = os;  // where edge c is read: v, or its group start in this line
        bool hit = has_c, miss = false;
        if (!has_c && ix.has_groups) {
          u32 pos = 0;
          if (group_start_in_line(dg, os, pos)) {
            if (STATS && pos != os) ++st_group;
            hit = pos != os && entry_bit(ds, pos);
            miss = !hit;
            at = pos;
          }
        }
        if (hit || miss) {  // else the full step of the next iteration decides
          if (STATS) ++st_stream;
          emitted = true;
          moved = true;
          if (hit) {
            v = cmap_of<Idx>(ix, c) + entry_base(ix, ds, bs, c) + entry_popc(ds, at);
            res = static_cast<u64>(v);
            if constexpr (KEY) {
              const u32 pr = walk_pred(dw, os, has_c, c);
              res |= static_cast<u64>(pr) << kPredShift;
              // a marked v's rank: its marks slot, read only for a marked v
              // (issued with the slots above, for every lane, it cost 4% on an
              // index of 5.1 GiB of lines and saved 5% of a 120 MiB one's)
              if (pr == kPredMarked && ix.line_ranks) {
                const uint4 dm = *line_slot(ix, bs, kSlotMarks);
                const u32 mr = dm.x + entry_popc(dm, os);
                if (mr <= kRankMax) res |= static_cast<u64>(mr + 1) << kRankShift;
              }
            }
          } else {  // the k-mer's last char made the miss: its predecessor was found
            res = kNotFound;
            prev = false;
            hint = kHintMiss;
            miss_at = p + k - 1;
          }
        }
      } else if (pass + 1 < passes && mode == kHash) {
        // a k-mer table probe, as in the full step
        const uint4 *group = ix.kmer_hash + 2 * hash_group();
        const uint4 h0 = group[0], h1 = group[1];
        u32 node = 0;
        const int pr = probe_group(h0, h1, hash_tag(), ix.kmer_hash_node_bits, node);
        moved = true;
        if (pr >= 0) {
          if (STATS) ++st_full;  // the first line stands in for the presearch line of a full search
          emitted = true;
          if (pr == 1) {
            v = static_cast<Idx>(node);
            prev = true;
            hint = kHintNone;
            res = static_cast<u64>(v);  // (KEY: the "look up" prediction is 0)
          } else {  // where the k-mer fails is unknown: check its last kp-mer next
            res = kNotFound;
            prev = false;
            hint = kHintMiss;
            miss_at = p + k - 1;
          }
        } else if (++j < kHashGroups) {
          if (STATS && (hash_group() & 3) == 0) ++st_hash;  // the next group opens a new line
        } else {  // a long probe run: the full search answers
          if (STATS) ++st_hash;  // (its presearch line is the search's own)
          pre_idx = window >> (64 - 2 * kp);
          mode = kPresearch;
        }
      }
      if (__ballot(moved) == 0) break;
    }
  }
  if constexpr (STATS) {
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 0), static_cast<u64>(st_full));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 1), static_cast<u64>(st_steps));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 2), static_cast<u64>(st_walk));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 3), static_cast<u64>(st_inv));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 4), static_cast<u64>(st_stream));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 5), static_cast<u64>(st_extra));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 6), static_cast<u64>(st_fb));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 7), static_cast<u64>(st_group));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 8), static_cast<u64>(st_tail));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 9), static_cast<u64>(st_tail_hit));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 10), static_cast<u64>(st_skip));
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + 11), static_cast<u64>(st_hash));
  }
}

// Key-k-mer move of the streaming search's results (move_to_key_kmer), one
// wave per work item, right to left in blocks of 64 k-mers. A found k-mer's
// result holds its node and, in bits 61-62, what the search learnt about the
// k-mer before it from that k-mer's node line (kPred*). The key of k-mer i is
// then: the key of k-mer i + 1 when that result says "same key" (k-mer i's
// node is unmarked and its walk's first step is k-mer i + 1's node), its own
// node when marked, kWalkStuck when its walk has no edge, else the table (or
// the walk, key_move); with a color table in the view, every chain end reads
// its key's color entry instead. Chains of "same key" are resolved with one ballot and a
// shuffle per block; the block's first k-mer hands its key and flag to the
// block on its left. Only chain ends read the key table, instead of every
// found k-mer. Work items never share a chain: each starts with a full search.
constexpr u32 kRunBlocks = 4;  // blocks of 64 k-mers whose chain ends are looked up together

// Eight waves per SIMD: the pass is a chain of dependent round trips per
// wave, so occupancy is its throughput (65 -> 63 registers for the eighth wave:
// C3's key runs 3.84 -> 3.05 ms).
template <typename Idx, bool STATS>
__global__ __launch_bounds__(kSearchThreads, 8) void k_key_runs(IndexView ix,
                                                            const ulonglong2 *__restrict__ chunks,
                                                            const u64 *__restrict__ chunk_ends,
                                                            u64 n_reads, u64 *__restrict__ out,
                                                            u64 *__restrict__ stats) {
  const u64 n_chunks = chunk_ends[n_reads - 1];
  const u32 lane = threadIdx.x & 63;
  const u64 n_waves = static_cast<u64>(gridDim.x) * (kSearchThreads / 64);
  const u64 wave = static_cast<u64>(blockIdx.x) * (kSearchThreads / 64) + threadIdx.x / 64;
  // narrow indexes: bits 32-60 may hold a rank (kRankShift)
  constexpr u64 kNodeMask = sizeof(Idx) == 4 ? 0xFFFFFFFFULL : (1ULL << kPredShift) - 1;
  u32 st_walk = 0;
  for (u64 ci = wave; ci < n_chunks; ci += n_waves) {
    const ulonglong2 item = chunks[ci];
    const u64 g0 = item.x;
    const u32 len = static_cast<u32>(item.y >> 48);
    // flag of the k-mer right of the block: its kPred* in bits 0-1, above
    // them 1 + the rank it carries (kRankShift) or 0
    u32 carry_flag = kPredLookup;
    u64 carry_key = 0;             // and its key
    // the block ending at hi_: lane 63 holds k-mer hi_ - 1, lane q k-mer
    // hi_ - 64 + q; none below hi_ - 64 or 0
    auto block_pos = [&](u32 hi_, bool &valid_) {
      const u32 lo_ = hi_ > 64 ? hi_ - 64 : 0;
      valid_ = hi_ > 0 && lane + hi_ >= lo_ + 64;
      return g0 + (valid_ ? hi_ + lane - 64 : 0);
    };
    // groups of kRunBlocks blocks, right to left: their results are loaded
    // together, then every chain end's lookup of the group is issued before
    // any is used, then the chains are resolved block by block
    for (u32 hi = len; hi > 0; hi = hi > 64 * kRunBlocks ? hi - 64 * kRunBlocks : 0) {
      u64 rv[kRunBlocks], pos[kRunBlocks], base[kRunBlocks];
      bool valid[kRunBlocks], brk[kRunBlocks];
      u32 flag[kRunBlocks];
#pragma unroll
      for (u32 b = 0; b < kRunBlocks; ++b) {
        pos[b] = block_pos(hi > 64 * b ? hi - 64 * b : 0, valid[b]);
        rv[b] = valid[b] ? out[pos[b]] : kNotFound;
      }
      u32 right_flag = carry_flag;
#pragma unroll
      for (u32 b = 0; b < kRunBlocks; ++b) {
        const bool found = (rv[b] >> 63) == 0;
        flag[b] = found ? (static_cast<u32>(rv[b] >> kPredShift) & 3u) |
                              (sizeof(Idx) == 4 ? (static_cast<u32>(rv[b] >> kRankShift) & ((1u << 29) - 1)) << 2
                                                : 0u)
                        : kPredLookup;
        const u64 node = rv[b] & kNodeMask;
        u32 succ = static_cast<u32>(__shfl_down(static_cast<int>(flag[b]), 1, 64));
        if (lane == 63) succ = right_flag;
        brk[b] = found && (succ & 3u) != kPredSameKey;
        base[b] = 0;
        if (brk[b]) {
          if ((succ & 3u) == kPredStuck) {
            base[b] = kWalkStuck;
          } else if (ix.color_table && (succ >> 2) != 0) {
            // a marked node's rank, from the search: its entry, or its mask's index
            base[b] = ix.color_table_kind == kColorTableMasks ? ix.key_masks[(succ >> 2) - 1] : (succ >> 2) - 1;
          } else if (ix.color_table) {  // the key's color entry (for a marked node, its own)
            base[b] = color_table_entry<Idx>(ix, static_cast<Idx>(node));
            if (STATS) ++st_walk;
          } else if ((succ & 3u) == kPredMarked) {
            base[b] = node;
          } else {
            u32 w = 0;
            base[b] = key_move<Idx>(ix, static_cast<Idx>(node), w);
            if (STATS) st_walk += w;
          }
        }
        right_flag = static_cast<u32>(__shfl(static_cast<int>(flag[b]), 0, 64));
      }
#pragma unroll
      for (u32 b = 0; b < kRunBlocks; ++b) {
        // a chain takes the key of the first break to its right, else the carry
        const u64 breaks = __ballot(brk[b]);
        const u64 right = lane == 63 ? 0 : breaks & (~0ULL << (lane + 1));
        const int src = right ? __ffsll(static_cast<long long>(right)) - 1 : 0;
        const u64 from = __shfl(base[b], src, 64);
        const u64 key = brk[b] ? base[b] : (right ? from : carry_key);
        if (valid[b] && (rv[b] >> 63) == 0) out[pos[b]] = key;
        carry_flag = static_cast<u32>(__shfl(static_cast<int>(flag[b]), 0, 64));
        carry_key = __shfl(key, 0, 64);
      }
    }
  }
  if constexpr (STATS) {
    if (st_walk) atomicAdd(reinterpret_cast<unsigned long long *>(stats + 2), static_cast<u64>(st_walk));
  }
}

// Absence checks pay when most kp-mers are in no k-mer of the index: with
// 4^kp >= 8 n at most about one kp-mer in eight has a non-empty interval.
inline bool absence_checks(const DeviceIndex &ix) {
  return ix.kp < ix.k && (1ULL << (2 * ix.kp)) >= 8 * ix.n;
}

inline u32 grid_for(u64 items, u32 per_block) {
  const u64 g = (items + per_block - 1) / per_block;
  return static_cast<u32>(std::max<u64>(1, std::min<u64>(g, kMaxGrid)));
}

template <typename F>
void dispatch_search(bool wide, bool key, F f) {
  if (wide) {
    if (key) f.template operator()<u64, true>();
    else f.template operator()<u64, false>();
  } else {
    if (key) f.template operator()<u32, true>();
    else f.template operator()<u32, false>();
  }
}

template <typename F>
void dispatch_search3(bool wide, bool key, bool stats, F f) {
  dispatch_search(wide, key, [&]<typename Idx, bool KEY>() {
    if (stats) f.template operator()<Idx, KEY, true>();
    else f.template operator()<Idx, KEY, false>();
  });
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free search, in 256-byte aligned parts (u64
// words): packed sequence (n/32 + 8 words), invalid masks (as many u32),
// inclusive k-mer ends per read, and the block sums of their scan.
// ---------------------------------------------------------------------------
namespace {
constexpr u64 align_words(u64 w) { return (w + 31) / 32 * 32; }
constexpr u64 kNaPackedWords = kNoAllocMaxChars / 32 + 8;
constexpr u64 kNaScanLen = kNoAllocMaxReads / kScanChunk + 2;
constexpr u64 kNaPacked = 0;
constexpr u64 kNaInvalid = kNaPacked + align_words(kNaPackedWords);
constexpr u64 kNaKmerEnds = kNaInvalid + align_words((kNaPackedWords + 1) / 2);
constexpr u64 kNaScanTmp = kNaKmerEnds + align_words(kNoAllocMaxReads);
constexpr u64 kNaWords = kNaScanTmp + align_words(2 * kNaScanLen);
}  // namespace
static_assert(kNaWords * 8 < (4u << 20), "the allocation-free scratch stays below 4 MiB");

static __device__ __align__(256) u64 scratch[kNaWords];

u64 *noalloc_scratch(int device) {
  static std::mutex mu;
  static u64 *ptrs[256] = {};
  if (device < 0 || device >= 256) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> g(mu);
  if (!ptrs[device]) {
    void *p = nullptr;
    SBWT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[device] = static_cast<u64 *>(p);
  }
  return ptrs[device];
}

void search_reads_noalloc(const DeviceIndex &index, bool move_to_key_kmer, const char *d_chars,
                          u64 n_chars, const u64 *d_read_ends, u64 n_reads, u64 *d_results,
                          hipStream_t stream) {
  if (n_chars > kNoAllocMaxChars || n_reads > kNoAllocMaxReads)
    throw std::runtime_error("batch exceeds the allocation-free search's maximum of 4194304 chars "
                             "and 262144 reads");
  if (move_to_key_kmer && !index.has_marks)
    throw std::runtime_error("move to key k-mer requested but the index has no key-k-mer marks");
  int dev = 0;
  SBWT_HIP(hipGetDevice(&dev));
  if (dev != index.device) throw std::runtime_error("the index's device is not the current device");
  if (n_reads == 0) return;
  u64 *base = noalloc_scratch(dev);
  u64 *packed = base + kNaPacked;
  u32 *invalid = reinterpret_cast<u32 *>(base + kNaInvalid);
  u64 *kmer_ends = base + kNaKmerEnds, *scan_tmp = base + kNaScanTmp;
  const IndexView ix = index.view();
  launch_pack(d_chars, n_chars, packed, invalid, stream);
  const u64 chunks = (n_reads + kScanChunk - 1) / kScanChunk;
  const u32 scan_grid = static_cast<u32>(chunks);
  k_scan_reduce<false><<<scan_grid, kScanThreads, 0, stream>>>(d_read_ends, n_reads, ix.k, n_chars, 1,
                                                               scan_tmp, kNaScanLen);
  k_scan_chunks<<<1, 1024, 0, stream>>>(scan_tmp, chunks, kNaScanLen);
  k_scan_write<false><<<scan_grid, kScanThreads, 0, stream>>>(d_read_ends, n_reads, ix.k, n_chars, 1,
                                                              scan_tmp, kNaScanLen, kmer_ends, nullptr);
  dispatch_search(index.wide, move_to_key_kmer, [&]<typename Idx, bool KEY>() {
    k_search_reads<Idx, KEY, false><<<grid_for(n_chars, kSearchThreads), kSearchThreads, 0, stream>>>(
        ix, packed, invalid, d_read_ends, kmer_ends, n_reads, n_chars, d_results, nullptr);
  });
  SBWT_HIP(hipGetLastError());
}

void launch_pack(const char *d_chars, u64 n_chars, u64 *d_packed, u32 *d_invalid, hipStream_t s) {
  const u64 n_words = (n_chars + 31) / 32 + 4;  // zero words of padding past the end
  const bool aligned = (reinterpret_cast<uintptr_t>(d_chars) % 16) == 0;
  k_pack<<<grid_for(n_words, 256), 256, 0, s>>>(reinterpret_cast<const u8 *>(d_chars), n_chars,
                                                n_words, aligned, d_packed, d_invalid);
  SBWT_HIP(hipGetLastError());
}

Searcher::Searcher(const DeviceIndex *index, u64 max_chars, u64 max_reads, bool move_to_key_kmer,
                   hipStream_t stream)
    : index_(index),
      stream_(stream),
      own_stream_(stream == nullptr),
      move_to_key_(move_to_key_kmer),
      max_chars_(max_chars),
      max_reads_(max_reads) {
  if (move_to_key_ && !index_->has_marks)
    throw std::runtime_error("move to key k-mer requested but the index has no key-k-mer marks");
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  const u64 words = max_chars_ / 32 + 8;
  SBWT_HIP(hipMalloc(&d_packed_, words * 8));
  SBWT_HIP(hipMalloc(&d_invalid_, words * 4));
  SBWT_HIP(hipMalloc(&d_kmer_ends_, std::max<u64>(max_reads_, 1) * 8));
  scan_tmp_len_ = max_reads_ / kScanChunk + 2;
  SBWT_HIP(hipMalloc(&d_scan_tmp_, 2 * scan_tmp_len_ * 8));  // k-mer and work-item block sums
  for (auto &e : ev_) SBWT_HIP(hipEventCreate(&e));
  SBWT_HIP(hipMalloc(&d_stats_, kStatsWords * 8));
  int dev = 0;
  SBWT_HIP(hipGetDevice(&dev));
  SBWT_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
  // work items: at most chars / chunk + reads + 1; an explicit chunk is at
  // least kMinChunkKmers, an automatic one gives no more items than lanes
  const u64 max_lanes = u64{kMaxBlocksPerCu} * static_cast<u64>(std::max(1, cus_)) * kSearchThreads;
  max_chunks_ = std::max<u64>(max_chars_ / kMinChunkKmers,
                              std::min<u64>(max_chars_ / kAutoMinChunkKmers + 1, max_lanes)) +
                max_reads_ + 1;
  SBWT_HIP(hipMalloc(&d_chunks_, max_chunks_ * 16));
  SBWT_HIP(hipMalloc(&d_chunk_counts_, std::max<u64>(max_reads_, 1) * 8 * 2));
  SBWT_HIP(hipMalloc(&d_counter_, 8));
  set_launch(kDefaultChunkKmers, kDefaultBlocksPerCu);
}

void Searcher::set_color_output(int mode) {
  if (mode < kColorOutputNone || mode > kColorOutputChains) throw std::runtime_error("unknown color output");
  if (mode != kColorOutputNone && (!move_to_key_ || !index_->color_table))
    throw std::runtime_error("color-table output needs the key-k-mer move and the index's color table");
  if (mode == kColorOutputChains && index_->wide)
    throw std::runtime_error("key-chain output needs an index of fewer than 2^32 nodes");
  color_output_ = mode;
}

IndexView Searcher::reads_view() const {
  IndexView v = index_->view();
  if (color_output_) {
    v.color_table = index_->color_table;
    v.color_table_kind = index_->color_table_kind;
    v.line_ranks = index_->color_ranks_in_lines ? 1u : 0u;
    v.key_masks = index_->key_color_masks;
  }
  return v;
}

void Searcher::set_launch(u32 chunk_kmers, u32 blocks_per_cu) {
  if (chunk_kmers < kMinChunkKmers || chunk_kmers > kMaxChunkKmers)
    throw std::runtime_error("chunk_kmers must be in [64, 65535]");
  if (blocks_per_cu < 1 || blocks_per_cu > kMaxBlocksPerCu)
    throw std::runtime_error("blocks_per_cu must be in [1, 8]");
  chunk_kmers_ = chunk_kmers;
  blocks_per_cu_ = blocks_per_cu;
  persistent_blocks_ = static_cast<u32>(std::max(1, cus_)) * blocks_per_cu;
}

u32 Searcher::stream_chunk(u64 n_chars) const {
  if (algorithm_ != kAlgoAuto && algorithm_ != kAlgoStreamingAdaptive) return chunk_kmers_;
  const u64 lanes = std::max<u64>(u64{persistent_blocks_} * kSearchThreads / lane_share_, 1);
  const u64 c = (n_chars + lanes - 1) / lanes;
  return static_cast<u32>(std::clamp<u64>(c, kAutoMinChunkKmers, chunk_kmers_));
}

LaunchShape Searcher::autotune(const char *d_chars, u64 n_chars, const u64 *d_read_ends, u64 n_reads,
                               u64 *d_results) {
  static constexpr u32 kChunks[] = {128, 256, 512, 1024};
  static constexpr u32 kBlocks[] = {2, 3, 4, 5, 6, 7, 8};
  const bool timing = timing_;
  const int algorithm = algorithm_;
  timing_ = true;
  algorithm_ = kAlgoStreaming;  // the shape only affects the streaming search
  LaunchShape best{chunk_kmers_, blocks_per_cu_, 0.0f};
  bool have = false;
  for (u32 ch : kChunks) {
    for (u32 bpc : kBlocks) {
      set_launch(ch, bpc);
      float ms = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {  // the second run is timed
        search_reads_device(d_chars, n_chars, d_read_ends, n_reads, d_results, nullptr);
        ms = last_timings().search_ms;
      }
      if (!have || ms < best.search_ms) best = {ch, bpc, ms};
      have = true;
    }
  }
  set_launch(best.chunk_kmers, best.blocks_per_cu);
  timing_ = timing;
  algorithm_ = algorithm;
  return best;
}

void Searcher::last_stats(u64 out[kStatsWords]) {
  SBWT_HIP(hipMemcpyAsync(out, d_stats_, kStatsWords * 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

Searcher::~Searcher() {
  for (auto &e : ev_) (void)hipEventDestroy(e);
  void *bufs[] = {d_packed_, d_invalid_, d_kmer_ends_, d_scan_tmp_, d_chars_, d_read_ends_,
                  d_results_, d_positions_, d_bits_, d_stats_, d_chunks_, d_chunk_counts_,
                  d_counter_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

void Searcher::ensure_staging(u64 chars, u64 reads, u64 positions) {
  if (chars > staged_chars_) {
    if (d_chars_) SBWT_HIP(hipFree(d_chars_));
    if (d_results_) SBWT_HIP(hipFree(d_results_));
    if (d_bits_) SBWT_HIP(hipFree(d_bits_));
    SBWT_HIP(hipMalloc(&d_chars_, chars + 64));
    SBWT_HIP(hipMalloc(&d_results_, (chars + 1) * 8));
    SBWT_HIP(hipMalloc(&d_bits_, (chars / 32 + 3) * 8));
    staged_chars_ = chars;
  }
  if (reads > staged_reads_) {
    if (d_read_ends_) SBWT_HIP(hipFree(d_read_ends_));
    SBWT_HIP(hipMalloc(&d_read_ends_, (reads + 1) * 8));
    staged_reads_ = reads;
  }
  if (positions > staged_positions_) {
    if (d_positions_) SBWT_HIP(hipFree(d_positions_));
    if (positions > staged_chars_) {
      if (d_results_) SBWT_HIP(hipFree(d_results_));
      SBWT_HIP(hipMalloc(&d_results_, (positions + 1) * 8));
      staged_chars_ = positions;
      if (d_chars_) SBWT_HIP(hipFree(d_chars_));
      SBWT_HIP(hipMalloc(&d_chars_, positions + 64));
      if (d_bits_) SBWT_HIP(hipFree(d_bits_));
      SBWT_HIP(hipMalloc(&d_bits_, (positions / 32 + 3) * 8));
    }
    SBWT_HIP(hipMalloc(&d_positions_, (positions + 1) * 8));
    staged_positions_ = positions;
  }
}

void Searcher::search_positions_device(const u64 *d_bit_seqs, const u64 *d_positions, u64 n,
                                       u64 *d_results) {
  if (n == 0) return;
  const IndexView ix = index_->view();
  if (timing_)
    for (int e = 0; e < 3; ++e) SBWT_HIP(hipEventRecord(ev_[e], stream_));
  dispatch_search(index_->wide, move_to_key_, [&]<typename Idx, bool KEY>() {
    k_search_positions<Idx, KEY><<<grid_for(n, kSearchThreads), kSearchThreads, 0, stream_>>>(
        ix, d_bit_seqs, d_positions, n, d_results);
  });
  SBWT_HIP(hipGetLastError());
  if (timing_) {
    SBWT_HIP(hipEventRecord(ev_[3], stream_));
    events_recorded_ = true;
    key_runs_timed_ = false;
  }
}

void Searcher::search_positions_host(const u64 *bit_seqs, u64 n_words, const u64 *positions, u64 n,
                                     u64 *results) {
  if (n == 0) return;
  for (u64 i = 0; i < n; ++i)
    if (positions[i] + index_->k > n_words * 32)
      throw std::runtime_error("k-mer position " + std::to_string(positions[i]) +
                               " reads past the end of bit_seqs");
  ensure_staging(std::max<u64>(n_words * 32, 32), 0, n);
  SBWT_HIP(hipMemsetAsync(d_bits_, 0, (staged_chars_ / 32 + 3) * 8, stream_));
  SBWT_HIP(hipMemcpyAsync(d_bits_, bit_seqs, n_words * 8, hipMemcpyHostToDevice, stream_));
  SBWT_HIP(hipMemcpyAsync(d_positions_, positions, n * 8, hipMemcpyHostToDevice, stream_));
  search_positions_device(d_bits_, d_positions_, n, d_results_);
  SBWT_HIP(hipMemcpyAsync(results, d_results_, n * 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

u64 Searcher::search_reads_device(const char *d_chars, u64 n_chars, const u64 *d_read_ends,
                                  u64 n_reads, u64 *d_results, u64 *n_kmers_host) {
  if (n_reads == 0) {
    if (n_kmers_host) *n_kmers_host = 0;
    return 0;
  }
  if (n_chars > max_chars_ || n_reads > max_reads_)
    throw std::runtime_error("batch exceeds the searcher's max_chars_per_batch / max_reads_per_batch");
  const IndexView ix = reads_view();
  if (timing_) SBWT_HIP(hipEventRecord(ev_[0], stream_));
  launch_pack(d_chars, n_chars, d_packed_, d_invalid_, stream_);
  if (timing_) SBWT_HIP(hipEventRecord(ev_[1], stream_));
  const bool streaming = !index_->groups_rejected &&
                         (algorithm_ == kAlgoStreaming ||
                          (algorithm_ == kAlgoAuto &&
                           n_chars >= (ix.kmer_hash ? kAutoStreamingCharsTable : kAutoStreamingChars)) ||
                          (algorithm_ == kAlgoStreamingAdaptive && n_chars >= kAdaptiveStreamingChars));
  // work items of the streaming search: (read, up to `chunk` k-mers)
  const u32 chunk = streaming ? stream_chunk(n_chars) : 1;
  if (streaming && n_chars / chunk + n_reads + 1 > max_chunks_)
    throw std::runtime_error("work-item table too small for the batch");
  u64 *d_item_ends = d_chunk_counts_ + max_reads_;
  const u64 chunks = (n_reads + kScanChunk - 1) / kScanChunk;
  const u32 scan_grid = static_cast<u32>(chunks);
  if (streaming) {
    k_scan_reduce<true><<<scan_grid, kScanThreads, 0, stream_>>>(d_read_ends, n_reads, ix.k, n_chars,
                                                                 chunk, d_scan_tmp_, scan_tmp_len_);
    k_scan_chunks<<<2, 1024, 0, stream_>>>(d_scan_tmp_, chunks, scan_tmp_len_);
    k_scan_write<true><<<scan_grid, kScanThreads, 0, stream_>>>(d_read_ends, n_reads, ix.k, n_chars,
                                                                chunk, d_scan_tmp_, scan_tmp_len_,
                                                                d_kmer_ends_, d_item_ends);
  } else {
    k_scan_reduce<false><<<scan_grid, kScanThreads, 0, stream_>>>(d_read_ends, n_reads, ix.k, n_chars,
                                                                  chunk, d_scan_tmp_, scan_tmp_len_);
    k_scan_chunks<<<1, 1024, 0, stream_>>>(d_scan_tmp_, chunks, scan_tmp_len_);
    k_scan_write<false><<<scan_grid, kScanThreads, 0, stream_>>>(d_read_ends, n_reads, ix.k, n_chars,
                                                                 chunk, d_scan_tmp_, scan_tmp_len_,
                                                                 d_kmer_ends_, nullptr);
  }
  SBWT_HIP(hipGetLastError());
  if (stats_) SBWT_HIP(hipMemsetAsync(d_stats_, 0, kStatsWords * 8, stream_));
  if (streaming) {
    // work-item table, then the persistent search
    k_chunk_fill<<<grid_for(n_reads, 256), 256, 0, stream_>>>(d_kmer_ends_, d_read_ends, n_reads,
                                                              d_item_ends, chunk, n_chars, d_chunks_);
    k_reset_counter<<<1, 1, 0, stream_>>>(d_counter_);
    SBWT_HIP(hipGetLastError());
    if (timing_) SBWT_HIP(hipEventRecord(ev_[2], stream_));
    // no more workgroups than there can be chunks for (a bound from the host:
    // reads plus chars per chunk)
    const u64 chunk_bound = n_reads + n_chars / chunk + 1;
    const u32 blocks = static_cast<u32>(
        std::min<u64>(persistent_blocks_, (chunk_bound + kSearchThreads - 1) / kSearchThreads));
    dispatch_search3(index_->wide, move_to_key_, stats_, [&]<typename Idx, bool KEY, bool STATS>() {
      k_search_stream<Idx, KEY, STATS><<<blocks, kSearchThreads, 0, stream_>>>(
          ix, d_packed_, d_invalid_, d_chunks_, d_item_ends, n_reads, d_counter_, d_results,
          d_stats_, absence_checks(*index_) ? 1u : 0u);
      // (key chains: the color searcher resolves them)
      if (KEY && color_output_ != kColorOutputChains) {
        if (timing_) SBWT_HIP(hipEventRecord(ev_[4], stream_));
        // one wave per work item
        k_key_runs<Idx, STATS><<<grid_for(chunk_bound, kSearchThreads / 64), kSearchThreads, 0,
                                 stream_>>>(ix, d_chunks_, d_item_ends, n_reads, d_results, d_stats_);
      }
    });
  } else {
    if (timing_) SBWT_HIP(hipEventRecord(ev_[2], stream_));
    // n_chars bounds the k-mer count; surplus workgroups leave at once. Key
    // chains: each found k-mer's node, every one a chain end of its own
    const bool key = move_to_key_ && color_output_ != kColorOutputChains;
    dispatch_search3(index_->wide, key, stats_, [&]<typename Idx, bool KEY, bool STATS>() {
      k_search_reads<Idx, KEY, STATS>
          <<<grid_for(n_chars, kSearchThreads), kSearchThreads, 0, stream_>>>(
              ix, d_packed_, d_invalid_, d_read_ends, d_kmer_ends_, n_reads, n_chars, d_results,
              d_stats_);
    });
  }
  SBWT_HIP(hipGetLastError());
  if (timing_) {
    SBWT_HIP(hipEventRecord(ev_[3], stream_));
    events_recorded_ = true;
    key_runs_timed_ = streaming && move_to_key_ && color_output_ != kColorOutputChains;
  }
  u64 nk = 0;
  if (n_kmers_host) {
    SBWT_HIP(hipMemcpyAsync(&nk, d_kmer_ends_ + (n_reads - 1), 8, hipMemcpyDeviceToHost, stream_));
    SBWT_HIP(hipStreamSynchronize(stream_));
    *n_kmers_host = nk;
  }
  return nk;
}

u64 Searcher::search_reads_host(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads,
                                u64 *results) {
  u64 nk = 0, st = 0;
  for (u64 r = 0; r < n_reads; ++r) {
    if (read_ends[r] < st || read_ends[r] > n_chars)
      throw std::runtime_error("read_ends must be non-decreasing and end at n_chars");
    const u64 len = read_ends[r] - st;
    if (len >= index_->k) nk += len - index_->k + 1;
    st = read_ends[r];
  }
  if (n_reads == 0) return 0;
  ensure_staging(std::max<u64>(n_chars, 1), n_reads, 0);
  if (n_chars) SBWT_HIP(hipMemcpyAsync(d_chars_, chars, n_chars, hipMemcpyHostToDevice, stream_));
  SBWT_HIP(hipMemcpyAsync(d_read_ends_, read_ends, n_reads * 8, hipMemcpyHostToDevice, stream_));
  search_reads_device(d_chars_, n_chars, d_read_ends_, n_reads, d_results_, nullptr);
  if (nk && results)
    SBWT_HIP(hipMemcpyAsync(results, d_results_, nk * 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  return nk;
}

void Searcher::search_reads_device_staged(const char *d_chars, u64 n_chars, const u64 *d_read_ends,
                                          u64 n_reads) {
  if (n_reads == 0) return;
  ensure_staging(std::max<u64>(n_chars, 1), 0, 0);
  search_reads_device(d_chars, n_chars, d_read_ends, n_reads, d_results_, nullptr);
}

SearchTimings Searcher::last_timings() {
  SearchTimings t;
  if (!events_recorded_) return t;
  SBWT_HIP(hipEventSynchronize(ev_[3]));
  SBWT_HIP(hipEventElapsedTime(&t.pack_ms, ev_[0], ev_[1]));
  SBWT_HIP(hipEventElapsedTime(&t.offsets_ms, ev_[1], ev_[2]));
  SBWT_HIP(hipEventElapsedTime(&t.search_ms, ev_[2], ev_[3]));
  if (key_runs_timed_) SBWT_HIP(hipEventElapsedTime(&t.key_runs_ms, ev_[4], ev_[3]));
  return t;
}

}  // namespace sbwt_amd
