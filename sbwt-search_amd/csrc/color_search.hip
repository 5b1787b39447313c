// Color search kernels for gfx950: per-read color counts in LDS, threshold,
// sorted-id compaction; plus the reference's warp-padded interface.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "color_search.hpp"
#include "device_fns.hpp"
#include "scan.hpp"

namespace sbwt_amd {

namespace {

constexpr u32 kWave = 64;
constexpr u64 kLdsColors = 1024;  // counters in LDS up to this many colors, else global
constexpr u64 kMaxGrid = 1u << 20;

template <typename T>
T *dev_upload(const T *host, u64 count) {
  T *d = nullptr;
  SBWT_HIP(hipMalloc(&d, std::max<u64>(count, 1) * sizeof(T)));
  if (count) SBWT_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

inline bool host_bit(const std::vector<u64> &w, u64 i) { return (w[i >> 6] >> (i & 63)) & 1ULL; }

inline u64 host_intv(const HostIntVector &v, u64 i) {
  const u64 bit = i * v.width, w = bit >> 6;
  const u32 off = static_cast<u32>(bit & 63);
  u64 x = v.words[w] >> off;
  if (off + v.width > 64) x |= v.words[w + 1] << (64 - off);
  return v.width == 64 ? x : (x & ((1ULL << v.width) - 1));
}

// ---------------------------------------------------------------------------
// Color set of a found (key) SBWT index: key-mark rank -> color_set_idx ->
// set table (ColorSearcher.cuh:99-131), the last two folded at load into one
// table by key rank (two dependent loads: the mark entry, the set entry): the
// payload's start (dense: first bit; sparse: first id) and length (dense:
// bits; sparse: ids).
// ---------------------------------------------------------------------------
struct SetRef {
  u64 start = 0;
  u32 len = 0;
  bool dense = false;
};


// Per-wave census of the color stage's algorithmic units, counted only when
// the kernel gets a census array: key-set entries read (one 128-byte line
// each), key-mark entries read (key-node input; one line each) and the
// 128-byte lines spanned by the decoded sets' payloads.
struct Census {
  bool on = false;
  u32 set_lines = 0, mark_lines = 0, payload_lines = 0;
};

constexpr u32 kCensusWords = 4;  // set lines, mark lines, payload lines, k-mers

// Windows of 64 result slots decoded together: their result loads, the set
// lookups of all their runs and the payload loads of those sets are each in
// flight at once, so a read of up to 256 k-mers costs about three memory
// round trips rather than three per window.
constexpr u32 kGroup = 4;
// Consecutive sparse ids per lane in the payload walk (loaded together
// before their counter adds), and dense 64-bit words per lane in flight.
constexpr u32 kPerLane = 8;
constexpr u32 kDenseUnroll = 2;
// Sparse ids of at most this width are walked in chunks of kPerLane ids of
// one set, each chunk one 16-byte load (8 * 12 bits + a 31-bit offset fit).
constexpr u32 kChunkWidth = 12;
static_assert(kPerLane * kChunkWidth + 31 <= 128 && kPerLane * kChunkWidth <= 96, "a chunk fits one load");
static_assert(kGroup * kWave < 0x10000, "a run length fits 16 bits of the chunk table");
// The wave's LDS table: 64 entries of 16 bytes.
constexpr u32 kTabWords = kWave * 4;

__device__ __forceinline__ u32 wave_sum(u32 x) {
  for (u32 o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__device__ __forceinline__ u32 wave_or(u32 x) {
  for (u32 o = kWave / 2; o > 0; o >>= 1) x |= __shfl_xor(x, o, kWave);
  return x;
}

__device__ __forceinline__ u64 read_lane64(u64 x, u32 l) {
  const u32 lo = __builtin_amdgcn_readlane(static_cast<u32>(x), l);
  const u32 hi = __builtin_amdgcn_readlane(static_cast<u32>(x >> 32), l);
  return static_cast<u64>(lo) | (static_cast<u64>(hi) << 32);
}

// 128-byte lines spanned by bits [b0, b1).
__device__ __forceinline__ u32 lines_spanned(u64 b0, u64 b1) {
  return b1 > b0 ? static_cast<u32>(((b1 - 1) >> 10) - (b0 >> 10) + 1) : 0u;
}

// Counters belong to one wave (LDS, or its own slice of a global array), so
// their atomics need only workgroup scope: device-scope ordering across the
// MI355X's eight XCD-local L2s would write back and invalidate L2.
__device__ __forceinline__ void cnt_add(u32 *p, u32 v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Units of the sets spread over the lanes: unit t of `total` (the sum of
// `units` over the lanes) belongs to the first lane whose inclusive prefix
// `incl` exceeds it; returns that lane and t's index in its set.
__device__ __forceinline__ u32 unit_owner(u32 incl, u32 units, u32 t, u32 &index) {
  u32 lo = 0, hi = kWave - 1;
#pragma unroll
  for (u32 step = 0; step < 6; ++step) {  // 64 lanes: six halvings
    const u32 mid = (lo + hi) >> 1;
    if (__shfl(incl, static_cast<int>(mid), kWave) > t) hi = mid;
    else lo = mid + 1;
  }
  index = t - (__shfl(incl, static_cast<int>(lo), kWave) - __shfl(units, static_cast<int>(lo), kWave));
  return lo;
}

// Inclusive prefix sum over the wave's lanes by DPP: shifts within each row
// of 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried by the
// row_bcast15 / row_bcast31 broadcasts into the rows above them.
__device__ __forceinline__ u32 wave_prefix(u32 v) {
  u32 x = v;
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x111, 0xF, 0xF, true));
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x112, 0xF, 0xF, true));
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x114, 0xF, 0xF, true));
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x118, 0xF, 0xF, true));
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x142, 0xA, 0xF, false));
  x += static_cast<u32>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x143, 0xC, 0xF, false));
  return x;
}

// Records that color c's counter word (64 colors) was written; only the
// touched words are scanned at the threshold (LDS counters: at most 16).
__device__ __forceinline__ void touch(u32 &touched, u32 c) { touched |= 1u << ((c >> 6) & 31); }

// The payloads of up to 64 decoded sets (one per lane), walked by the whole
// wave, adding each set's run length to the counter of every color it holds
// (the per-color ballot/popcount of ColorSearcher.cuh:175-205 summed over
// the read by ColorPostProcessor.cuh:21-43). `tab` is the wave's LDS table
// of 64 entries.
//  - dense sets: their 64-bit bitmap words are loaded one per lane, two per
//    lane in flight; then the wave steps through the words, each broadcast
//    from its lane (readlane), and lane i adds the run to color 64 u + i when
//    its bit is set: no bit loop and no two lanes on one counter;
//  - sparse sets of ids up to kChunkWidth bits: each set is cut into chunks
//    of kPerLane ids, one chunk per lane (a search over the lanes' chunk
//    counts finds its set); the chunk is one 16-byte load, shifted to bit 0,
//    and id k is the width bits at k * width;
//  - wider sparse ids: the ids of all sets, concatenated, are cut into ranges
//    of kPerLane consecutive ids, one range per lane; a lane finds the set of
//    its first id once and steps on through the table of set ranges. All
//    loads of a range are issued before any id is used (one memory round trip
//    per range), then the ids are extracted and added.
__device__ __forceinline__ void wave_payload(const ColorView &cv, const SetRef &set, u32 units,
                                             u32 run, u32 lane, u32 *cnt, uint4 *tab,
                                             u32 &touched) {
  const u32 nc = cv.num_colors;
  const u32 d_units = set.dense ? units : 0u, s_units = set.dense ? 0u : units;
  if (__ballot(d_units != 0)) {
    const u32 d_incl = wave_prefix(d_units);
    const u32 d_total = __builtin_amdgcn_readlane(d_incl, kWave - 1);
    for (u32 t0 = 0; t0 < d_total; t0 += kDenseUnroll * kWave) {
      // loads first (both words of every unit, all in flight), then the bits
      u64 x0[kDenseUnroll], x1[kDenseUnroll];
      u32 rq[kDenseUnroll], uq[kDenseUnroll], off[kDenseUnroll], take[kDenseUnroll];
#pragma unroll
      for (u32 q = 0; q < kDenseUnroll; ++q) {
        const u32 t = min(t0 + q * kWave + lane, d_total - 1);
        u32 u = 0;
        const u32 owner = unit_owner(d_incl, d_units, t, u);
        rq[q] = __shfl(run, static_cast<int>(owner), kWave);
        const u32 o_lo = __shfl(static_cast<u32>(set.start), static_cast<int>(owner), kWave);
        const u32 o_hi = __shfl(static_cast<u32>(set.start >> 32), static_cast<int>(owner), kWave);
        const u32 o_len = __shfl(set.len, static_cast<int>(owner), kWave);
        uq[q] = u;
        // 64 bits of the bitmap: colors 64u .. 64u + 63
        const u64 bit = (static_cast<u64>(o_lo) | (static_cast<u64>(o_hi) << 32)) + 64 * static_cast<u64>(u);
        off[q] = static_cast<u32>(bit & 63);
        take[q] = t0 + q * kWave + lane < d_total ? min(64u, o_len - 64 * u) : 0u;
        x0[q] = cv.dense_arrays[bit >> 6];
        x1[q] = cv.dense_arrays[(bit >> 6) + 1];
      }
#pragma unroll
      for (u32 q = 0; q < kDenseUnroll; ++q) {
        if (t0 + q * kWave >= d_total) continue;  // wave-uniform
        u64 x = off[q] ? ((x0[q] >> off[q]) | (x1[q] << (64 - off[q]))) : x0[q];
        if (take[q] < 64) x &= (1ULL << take[q]) - 1;
        const u32 n_here = min(kWave, d_total - t0 - q * kWave);
        for (u32 l = 0; l < n_here; ++l) {
          const u32 xl = __builtin_amdgcn_readlane(static_cast<u32>(x), l);
          const u32 xh = __builtin_amdgcn_readlane(static_cast<u32>(x >> 32), l);
          if ((xl | xh) == 0) continue;
          const u32 r = __builtin_amdgcn_readlane(rq[q], l);
          const u32 u = __builtin_amdgcn_readlane(uq[q], l);
          const u32 bitv = lane < 32 ? (xl >> lane) : (xh >> (lane - 32));
          const u32 c = 64 * u + lane;
          if ((bitv & 1u) && c < nc) {
            cnt_add(cnt + c, r);
            touch(touched, c);
          }
        }
      }
    }
  }
  if (__ballot(s_units != 0) == 0) return;
  const u32 width = cv.sparse_width;
  if (width <= kChunkWidth && __ballot(s_units > 0xFFFFu) == 0) {
    // chunks of up to kPerLane ids of one set, one chunk per lane: the
    // chunk's ids lie in 16 bytes (kPerLane * width + 31 <= 128 bits), one
    // load; shifted to bit 0, id k sits at bit k * width of 96 bits
    const u32 c_units = (s_units + kPerLane - 1) / kPerLane;
    const u32 c_incl = wave_prefix(c_units);
    const u32 c_total = __builtin_amdgcn_readlane(c_incl, kWave - 1);
    tab[lane] = make_uint4(c_incl, static_cast<u32>(set.start), static_cast<u32>(set.start >> 32),
                           run | (s_units << 16));
    wave_lds_fence();
    const u32 *sparse32 = reinterpret_cast<const u32 *>(cv.sparse_arrays);
    const u32 wmask = (1u << width) - 1;
    for (u32 b0 = 0; b0 < c_total; b0 += kWave) {
      u32 u = 0;
      const u32 o = unit_owner(c_incl, c_units, min(b0 + lane, c_total - 1), u);
      const uint4 e = tab[o];
      const u32 r = e.w & 0xFFFFu;
      const u32 n = b0 + lane < c_total ? min(kPerLane, (e.w >> 16) - kPerLane * u) : 0u;
      const u64 bit = ((static_cast<u64>(e.y) | (static_cast<u64>(e.z) << 32)) + kPerLane * u) * width;
      const uint4 w = *reinterpret_cast<const uint4 *>(sparse32 + (bit >> 5));
      const u32 sh = static_cast<u32>(bit) & 31u;
      const u32 a0 = __builtin_amdgcn_alignbit(w.y, w.x, sh), a1 = __builtin_amdgcn_alignbit(w.z, w.y, sh),
                a2 = __builtin_amdgcn_alignbit(w.w, w.z, sh);
      const u64 lo = static_cast<u64>(a0) | (static_cast<u64>(a1) << 32);  // bits 0-63
      const u64 hi = static_cast<u64>(a1) | (static_cast<u64>(a2) << 32);  // bits 32-95
#pragma unroll
      for (u32 k = 0; k < kPerLane; ++k) {
        const u32 pos = k * width;  // wave-uniform
        const u32 id = static_cast<u32>(pos + width <= 64 ? lo >> pos : hi >> (pos - 32)) & wmask;
        if (k < n && id < nc) {
          cnt_add(cnt + id, r);
          touch(touched, id);
        }
      }
    }
    wave_lds_fence();  // the table is written again by the next chunk
    return;
  }
  const u32 s_incl = wave_prefix(s_units);
  const u32 s_total = __builtin_amdgcn_readlane(s_incl, kWave - 1);
  // the set ranges: {end of the set's ids in the concatenation, start lo,
  // start hi, run}, lane order
  tab[lane] = make_uint4(s_incl, static_cast<u32>(set.start), static_cast<u32>(set.start >> 32), run);
  wave_lds_fence();
  for (u32 b0 = 0; b0 < s_total; b0 += kPerLane * kWave) {
    const u32 t = b0 + lane * kPerLane;  // this lane's first id
    const u32 t_first = min(t, s_total - 1);
    u32 u = 0;
    u32 o = unit_owner(s_incl, s_units, t_first, u);
    uint4 e = tab[o];
    u32 excl = t_first - u;  // the set's first id in the concatenation
    if (width <= 32) {
      // phase 1: a bit cursor over the lane's ids (one element on per id; the
      // next set's first element when the range crosses into it), each id's
      // two dwords loaded (all in flight) and {bit offset, run, valid} kept
      uint2 w[kPerLane];
      u32 meta[kPerLane];
      u64 bit = ((static_cast<u64>(e.y) | (static_cast<u64>(e.z) << 32)) + (t_first - excl)) * width;
      u32 end = e.x;                    // the current set's end in the concatenation
      u32 rs = min(e.w, 1023u) << 5;    // its run, placed for meta
      const u32 *sparse32 = reinterpret_cast<const u32 *>(cv.sparse_arrays);
#pragma unroll
      for (u32 k = 0; k < kPerLane; ++k) {
        const u32 tk = t + k;
        if (tk >= end && tk < s_total) {
          // on to the next set with ids (the ends rise); tk is its first id
          do {
            e = tab[++o];
          } while (e.x <= tk && o < kWave - 1);
          bit = (static_cast<u64>(e.y) | (static_cast<u64>(e.z) << 32)) * width;
          end = e.x;
          rs = min(e.w, 1023u) << 5;
        }
        w[k] = *reinterpret_cast<const uint2 *>(sparse32 + (bit >> 5));
        meta[k] = (static_cast<u32>(bit) & 31u) | rs | (tk < s_total ? 1u << 15 : 0u);
        if (tk + 1 < s_total) bit += width;  // past the last id the cursor stays in the arrays
      }
      // phase 2: the ids and their counter adds
      const u32 wmask = width == 32 ? ~0u : ((1u << width) - 1);
#pragma unroll
      for (u32 k = 0; k < kPerLane; ++k) {
        const u32 id = __builtin_amdgcn_alignbit(w[k].y, w[k].x, meta[k] & 31u) & wmask;
        if ((meta[k] >> 15) && id < nc) {
          cnt_add(cnt + id, (meta[k] >> 5) & 1023u);
          touch(touched, id);
        }
      }
    } else {
      for (u32 k = 0; k < kPerLane; ++k) {
        const u32 tk = t + k;
        if (tk >= s_total) break;
        while (tk >= e.x && o < kWave - 1) {
          excl = e.x;
          e = tab[++o];
        }
        const u64 id = intv_get(cv.sparse_arrays, width,
                                (static_cast<u64>(e.y) | (static_cast<u64>(e.z) << 32)) + (tk - excl));
        if (id < nc) {
          cnt_add(cnt + id, e.w);
          touch(touched, static_cast<u32>(id));
        }
      }
    }
  }
  wave_lds_fence();  // the table is written again by the next chunk
}

// Inputs of the color stage: each k-mer's key node (the reference's index
// output), or its key's entry of the index's color table: a color mask (the
// counts then come from ballots, with no lookups) or the key's rank; or the
// key chains of either (a searcher's kColorOutputChains), resolved into those
// entries here.
constexpr int kInKeys = 0, kInMasks = 1, kInRanks = 2, kInChainMasks = 3, kInChainRanks = 4;
// The entries an input is counted as (after resolving its chains).
constexpr int entry_input(int in) {
  return in == kInChainMasks ? kInMasks : (in == kInChainRanks ? kInRanks : in);
}
constexpr bool chain_input(int in) { return in == kInChainMasks || in == kInChainRanks; }

// Flag of a key-chain result for the k-mer before it: kPred* in bits 0-1, 1 +
// the rank of that k-mer's (marked) node above them or 0; a k-mer not found
// says nothing (kPredLookup).
__device__ __forceinline__ u32 chain_flag(u64 v) {
  return (v >> 63) ? kPredLookup
                   : (static_cast<u32>(v >> kPredShift) & 3u) |
                         ((static_cast<u32>(v >> kRankShift) & ((1u << 29) - 1)) << 2);
}

// The key chains of one group (kGroup windows of 64 results), right to left,
// as k_key_runs resolves them: a found k-mer whose successor's flag does not
// say "same key" ends a chain and takes its key's entry (a marked node's rank
// from the flag: the rank itself, or the index of its mask; stuck: the stuck
// sentinel; else the node's color-table entry); every other found k-mer takes
// the entry of the first chain end to its right. carry_flag / carry_key: the
// flag of the k-mer after the group and that k-mer's entry (in: from the
// group to the right; out: for the group to the left). Lines read: cs.
template <bool MASKS>
__device__ __forceinline__ void resolve_chains(const ColorView &cv, u64 (&v)[kGroup], u32 lane,
                                               u32 &carry_flag, u64 &carry_key, Census &cs) {
  // entries as the table holds them: masks, or u32 ranks with kNoKeyRank for stuck
  using Entry = std::conditional_t<MASKS, u64, u32>;
  u32 flag[kGroup];
  bool brk[kGroup];
  Entry base[kGroup];
#pragma unroll
  for (u32 g = 0; g < kGroup; ++g) flag[g] = chain_flag(v[g]);
  u32 right = carry_flag;  // flag of the slot after window g's last
#pragma unroll
  for (int g = kGroup - 1; g >= 0; --g) {
    u32 succ = static_cast<u32>(__shfl_down(static_cast<int>(flag[g]), 1, kWave));
    if (lane == kWave - 1) succ = right;
    right = static_cast<u32>(__shfl(static_cast<int>(flag[g]), 0, kWave));
    brk[g] = (v[g] >> 63) == 0 && (succ & 3u) != kPredSameKey;
    base[g] = 0;
    if (brk[g]) {
      if ((succ & 3u) == kPredStuck) {
        base[g] = MASKS ? static_cast<Entry>(kWalkStuck) : static_cast<Entry>(kNoKeyRank);
      } else if ((succ >> 2) != 0) {
        if constexpr (MASKS) {
          base[g] = cv.chain_key_masks[(succ >> 2) - 1];
          if (cs.on) ++cs.mark_lines;
        } else {
          base[g] = (succ >> 2) - 1;
        }
      } else {
        base[g] = static_cast<const Entry *>(cv.chain_table)[static_cast<u32>(v[g])];
        if (cs.on) ++cs.mark_lines;
      }
    }
  }
  Entry ckey = static_cast<Entry>(carry_key);
#pragma unroll
  for (int g = kGroup - 1; g >= 0; --g) {
    const u64 breaks = __ballot(brk[g]);
    const u64 rgt = lane == kWave - 1 ? 0 : breaks & (~0ULL << (lane + 1));
    const int src = rgt ? __ffsll(static_cast<long long>(rgt)) - 1 : 0;
    const Entry from = __shfl(base[g], src, kWave);
    const Entry key = brk[g] ? base[g] : (rgt ? from : ckey);
    if ((v[g] >> 63) == 0) v[g] = MASKS || key != kNoKeyRank ? static_cast<u64>(key) : kWalkStuck;
    ckey = __shfl(key, 0, kWave);
  }
  carry_key = ckey;
  carry_flag = static_cast<u32>(__shfl(static_cast<int>(flag[0]), 0, kWave));
}

// Up to kGroup windows of 64 result slots from slot `base` of a read ending
// at `ke`: counts found / not-found / invalid slots and adds every found
// k-mer's colors to the wave's counters.
// Key nodes and ranks: consecutive equal found results (consecutive k-mers
// of a read usually reach the same key k-mer) form runs, also across the
// windows; the first slot of each run (its leader) writes {result, run} to
// the wave's LDS table at its rank among the leaders, 64 leaders at a time,
// lane j takes entry j and decodes its set (the key-mark entry then the key
// set, or the key set alone) and the wave walks the payloads, adding the
// run's length. Masks: color c gains the number of slots whose mask holds
// it, lane c adding it.
template <int IN>
__device__ __forceinline__ void group_add(const ColorView &cv, const u64 *__restrict__ results,
                                          u64 base, u64 ke, u32 lane, u32 *cnt, uint4 *tab,
                                          u32 &f, u32 &nf, u32 &inv, u32 &touched, Census &cs,
                                          u32 &carry_flag, u64 &carry_key) {
  constexpr int E = entry_input(IN);
  u64 v[kGroup];
  bool fd[kGroup];
#pragma unroll
  for (u32 g = 0; g < kGroup; ++g) {
    const u64 i = base + g * kWave + lane;
    v[g] = i < ke ? results[i] : kInvalid;
  }
  if constexpr (chain_input(IN)) resolve_chains<IN == kInChainMasks>(cv, v, lane, carry_flag, carry_key, cs);
#pragma unroll
  for (u32 g = 0; g < kGroup; ++g) {
    const bool in = base + g * kWave + lane < ke;
    fd[g] = in && (E == kInKeys ? v[g] < cv.num_nodes : (v[g] >> 63) == 0);
    f += static_cast<u32>(__popcll(__ballot(fd[g])));
    inv += static_cast<u32>(__popcll(__ballot(in && v[g] == kInvalid)));
    nf += static_cast<u32>(__popcll(__ballot(in && !fd[g] && v[g] != kInvalid)));
  }
  if constexpr (E == kInMasks) {
    u32 mine = 0;
#pragma unroll
    for (u32 g = 0; g < kGroup; ++g) {
      const u64 m = fd[g] ? v[g] : 0;
      if (__ballot(m != 0) == 0) continue;
      for (u32 c = 0; c < cv.num_colors; ++c) {
        const u32 n = static_cast<u32>(__popcll(__ballot((m >> c) & 1)));
        if (lane == c) mine += n;
      }
    }
    if (mine) {
      cnt_add(cnt + lane, mine);
      touch(touched, lane);
    }
    return;
  }
  // same_m[g]: slots equal to the slot before them (lane 0 of a window
  // follows lane 63 of the window before); lead_m[g]: the runs' leaders
  u64 same_m[kGroup], lead_m[kGroup];
#pragma unroll
  for (u32 g = 0; g < kGroup; ++g) {
    const u64 last = g ? read_lane64(v[g - 1], kWave - 1) : 0;
    u64 prev = __shfl_up(v[g], 1, kWave);
    if (lane == 0) prev = g ? last : ~v[g];
    const bool same = fd[g] && prev == v[g];
    same_m[g] = __ballot(same);
    lead_m[g] = __ballot(fd[g] && !same);
  }
  // cont[g]: slots from window g's start that continue the run before it;
  // cum[g]: leaders before window g
  u32 cont[kGroup + 1], cum[kGroup + 1];
  cont[kGroup] = 0;
#pragma unroll
  for (int g = kGroup - 1; g >= 0; --g) {
    const u32 ones = same_m[g] == ~0ULL ? 64u : static_cast<u32>(__builtin_ctzll(~same_m[g]));
    cont[g] = ones + (ones == 64 ? cont[g + 1] : 0u);
  }
  cum[0] = 0;
#pragma unroll
  for (u32 g = 0; g < kGroup; ++g) cum[g + 1] = cum[g] + static_cast<u32>(__popcll(lead_m[g]));
  const u32 n_lead = cum[kGroup];
  for (u32 j0 = 0; j0 < n_lead; j0 += kWave) {
#pragma unroll
    for (u32 g = 0; g < kGroup; ++g) {
      if (cum[g + 1] <= j0 || cum[g] >= j0 + kWave) continue;  // no leader of this chunk here
      const u64 lm = lead_m[g];
      const u32 below = __builtin_amdgcn_mbcnt_hi(static_cast<u32>(lm >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(static_cast<u32>(lm), 0u));
      const u32 rank = cum[g] + below;
      if (((lm >> lane) & 1) && rank >= j0 && rank < j0 + kWave) {
        // the run: the leader, the equal slots after it in its window and,
        // when they reach the window's end, the slots continuing it later
        const u64 rest = lane == kWave - 1 ? 0 : (same_m[g] >> (lane + 1));
        const u32 ones = static_cast<u32>(__builtin_ctzll(~rest));
        const u32 run = 1 + ones + (ones == kWave - 1 - lane ? cont[g + 1] : 0u);
        tab[rank - j0] = make_uint4(static_cast<u32>(v[g]), static_cast<u32>(v[g] >> 32), run, 0u);
      }
    }
    wave_lds_fence();
    const bool has = j0 + lane < n_lead;
    const uint4 le = has ? tab[lane] : make_uint4(0u, 0u, 0u, 0u);
    wave_lds_fence();
    const u64 x = static_cast<u64>(le.x) | (static_cast<u64>(le.y) << 32);
    SetRef set;
    u32 units = 0;
    if (has) {
      const u64 rank = E == kInKeys ? rank_marks(cv.key_marks, x) : x;
      const uint4 e = cv.key_sets[rank];
      set.start = static_cast<u64>(e.x) | (static_cast<u64>(e.y) << 32);
      set.len = e.z;
      set.dense = (e.w & 1u) != 0;
      units = set.dense ? (set.len + 63) / 64 : set.len;
      if (cs.on) {
        ++cs.set_lines;
        cs.mark_lines += E == kInKeys ? 1u : 0u;
        const u64 w = set.dense ? 1 : cv.sparse_width;
        cs.payload_lines += lines_spanned(set.start * w, (set.start + set.len) * w);
      }
    }
    wave_payload(cv, set, units, le.z, lane, cnt, tab, touched);
  }
}

// The k-mer results [ks, ke) of one read (whose k-mers end at re) into the
// wave's counters. Key chains run right to left, group by group; a chain
// that continues past ke (into another work item) ends at ke - 1 with its
// node's color-table entry.
template <int IN>
__device__ __forceinline__ void read_add(const ColorView &cv, const u64 *__restrict__ results, u64 ks,
                                         u64 ke, u64 re, u32 lane, u32 *cnt, uint4 *tab, u32 &f,
                                         u32 &nf, u32 &inv, u32 &touched, Census &cs) {
  u32 carry_flag = kPredLookup;
  u64 carry_key = 0;
  if constexpr (chain_input(IN)) {
    if (ke <= ks) return;
    if (ke < re) {
      const u32 nf_ = chain_flag(results[ke]);
      carry_flag = (nf_ & 3u) == kPredSameKey ? kPredLookup : nf_;
    }
    constexpr u64 kStep = kGroup * kWave;
    for (u64 gi = (ke - ks - 1) / kStep + 1; gi-- > 0;)
      group_add<IN>(cv, results, ks + gi * kStep, ke, lane, cnt, tab, f, nf, inv, touched, cs,
                    carry_flag, carry_key);
  } else {
    for (u64 base = ks; base < ke; base += kGroup * kWave)
      group_add<IN>(cv, results, base, ke, lane, cnt, tab, f, nf, inv, touched, cs, carry_flag,
                    carry_key);
  }
}

// ---------------------------------------------------------------------------
// Per-wave color counters: nc u32 in LDS when nc <= kLdsColors (16 KiB per
// workgroup of four waves at most, so LDS never limits occupancy), else a
// slice of a global array per wave of a bounded grid (any number of colors;
// the payload atomics go to L2). LDS selects the storage at compile time.
// Global counters are read past L1 (workgroup scope: only their wave touches
// them) and are zero between reads.
// ---------------------------------------------------------------------------
template <bool LDS>
__device__ __forceinline__ u32 cnt_get(const u32 *p) {
  if constexpr (LDS) {
    return *p;
  } else {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// The wave's counter updates are complete and visible to its own reads.
template <bool LDS>
__device__ __forceinline__ void cnt_fence() {
  if constexpr (LDS) {
    wave_lds_fence();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  }
}

// The workgroup's LDS: each wave's table (kTabWords u32), then with LDS
// counters each wave's nc counters.
__device__ __forceinline__ uint4 *wave_table(u32 *lds) {
  return reinterpret_cast<uint4 *>(lds) + (threadIdx.x / kWave) * kWave;
}

template <bool LDS>
__device__ __forceinline__ u32 *wave_counters(u32 *lds, u32 *global_cnt, u32 nc) {
  const u32 wave = threadIdx.x / kWave;
  if constexpr (LDS) {
    return lds + (blockDim.x / kWave) * kTabWords + static_cast<u64>(wave) * nc;
  } else {
    return global_cnt + (static_cast<u64>(blockIdx.x) * (blockDim.x / kWave) + wave) * nc;
  }
}

// The wave's census into census[0..3] (set lines, mark lines, payload
// lines, k-mers), one atomic each.
__device__ __forceinline__ void census_flush(const Census &cs, u64 kmers, u32 lane,
                                             u64 *census) {
  if (!cs.on) return;
  const u32 sl = wave_sum(cs.set_lines), ml = wave_sum(cs.mark_lines);
  const u32 pl = wave_sum(cs.payload_lines);
  if (lane == 0) {
    atomicAdd(reinterpret_cast<unsigned long long *>(census), static_cast<unsigned long long>(sl));
    atomicAdd(reinterpret_cast<unsigned long long *>(census + 1), static_cast<unsigned long long>(ml));
    atomicAdd(reinterpret_cast<unsigned long long *>(census + 2), static_cast<unsigned long long>(pl));
    atomicAdd(reinterpret_cast<unsigned long long *>(census + 3), static_cast<unsigned long long>(kmers));
  }
}

struct KernelParams {
  double threshold;
  u32 include_not_found;
  u32 include_invalid;
  u32 first_read_quirk;
};

// Outputs of the threshold stage per read: the hit bitmap (words per read),
// the number of hits, optional {found, not found, invalid} and the quirk-2
// inputs (colored flag, threshold of a read that found nothing).
struct ReadOut {
  u64 *bitmap;
  u64 *nout;
  u32 *stats;
  u64 *blank_min;  // null: no quirk-2 pass follows
  u64 *colored;
};

// Read r's colors with count >= ceil(seq_size * tau)
// (ContinuousColorResultsPrinter.hpp:235-263) into its bitmap row, from the
// counters cnt (zeroed on the way); f/nf/inv are the read's whole counts.
// With LDS counters only the words in `touched` (this lane's record of the
// counter words it wrote; at most 16 words) are scanned: every other counter
// is zero, so no color there reaches a positive threshold.
template <bool LDS>
__device__ __forceinline__ void finish_read(u32 *cnt, u32 nc, u64 r, u32 f, u32 nf, u32 inv,
                                            const KernelParams &p, const ReadOut &o, u32 lane,
                                            u32 touched) {
  u64 F = f, NF = nf, INV = inv;
  if (p.first_read_quirk && r == 0) {
    F -= 1;
    NF -= 1;
    INV -= 1;
  }
  const u64 seq = F + (p.include_not_found ? NF : 0) + (p.include_invalid ? INV : 0);
  const u64 min_hits = static_cast<u64>(ceil(static_cast<double>(seq) * p.threshold));
  const u32 words = (nc + 63) / 64;
  u64 total = 0;
  if constexpr (LDS) {
    u64 mine = 0;  // lane w: bitmap word w
    for (u32 tw = wave_or(touched); tw; tw &= tw - 1) {
      const u32 w = static_cast<u32>(__builtin_ctz(tw));
      const u32 c = w * 64 + lane;
      const u32 v = c < nc ? cnt[c] : 0u;
      const u64 m = __ballot(min_hits > 0 && c < nc && v >= min_hits);
      if (lane == w) mine = m;
      total += __popcll(m);
      if (v) cnt[c] = 0;
    }
    if (lane < words) o.bitmap[r * words + lane] = mine;
  } else {
    for (u32 w = 0; w < words; ++w) {
      const u32 c = w * 64 + lane;
      const u32 v = c < nc ? cnt_get<LDS>(cnt + c) : 0u;
      const bool hit = min_hits > 0 && c < nc && v >= min_hits;
      const u64 m = __ballot(hit);
      if (lane == 0) o.bitmap[r * words + w] = m;
      total += __popcll(m);
      if (c < nc && v) cnt[c] = 0;
    }
  }
  if (lane == 0) {
    o.nout[r] = total;
    if (o.blank_min) {
      o.colored[r] = f > 0 ? 1 : 0;
      o.blank_min[r] = f == 0 ? min_hits : 0;
    }
    if (o.stats) {
      o.stats[3 * r] = f;
      o.stats[3 * r + 1] = nf;
      o.stats[3 * r + 2] = inv;
    }
  }
}

// Fused color search, one wave per read (grid-stride): the plain form.
// COUNTS writes the raw u32 counts; otherwise finish_read thresholds them.
// Key-node input holds five waves per SIMD under a register cap (98 -> 96:
// C3's key-node color kernel 14.6 -> 13.9 ms); a cap on the other inputs spills.
template <bool COUNTS, bool LDS, int IN>
__global__ __launch_bounds__(256, (IN == kInKeys ? 5 : 1)) void k_color_reads(ColorView cv, const u64 *__restrict__ results,
                                                      const u64 *__restrict__ kmer_ends, u64 n_reads,
                                                      KernelParams p, ReadOut o,
                                                      u32 *__restrict__ counts_out,
                                                      u32 *__restrict__ global_cnt,
                                                      u64 *__restrict__ census) {
  extern __shared__ __align__(16) u32 lds[];
  uint4 *tab = wave_table(lds);
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const u32 waves = blockDim.x / kWave;
  const u32 nc = cv.num_colors;
  u32 *cnt = wave_counters<LDS>(lds, global_cnt, nc);
  if constexpr (LDS) {
    for (u32 c = lane; c < nc; c += kWave) cnt[c] = 0;
  }
  cnt_fence<LDS>();
  Census cs;
  cs.on = census != nullptr;
  u64 kmers = 0;
  for (u64 r = static_cast<u64>(blockIdx.x) * waves + wave; r < n_reads;
       r += static_cast<u64>(gridDim.x) * waves) {
    const u64 ks = r ? kmer_ends[r - 1] : 0, ke = kmer_ends[r];
    u32 f = 0, nf = 0, inv = 0, touched = 0;
    read_add<IN>(cv, results, ks, ke, ke, lane, cnt, tab, f, nf, inv, touched, cs);
    kmers += ke - ks;
    cnt_fence<LDS>();
    if constexpr (COUNTS) {
      for (u32 c = lane; c < nc; c += kWave) {
        const u32 v = cnt_get<LDS>(cnt + c);
        counts_out[r * nc + c] = v;
        if (v) cnt[c] = 0;
      }
    } else {
      finish_read<LDS>(cnt, nc, r, f, nf, inv, p, o, lane, touched);
    }
    cnt_fence<LDS>();
  }
  census_flush(cs, kmers, lane, census);
}

// Work items of the batched form: read r is split into
// items[r] = max(1, ceil(k-mers / item_kmers)) items of consecutive k-mers;
// multi[r] = 1 when it has more than one (its partial counts meet in a slot).
__global__ __launch_bounds__(256) void k_color_item_counts(const u64 *__restrict__ kmer_ends,
                                                            u64 n_reads, u32 item_kmers,
                                                            u64 *__restrict__ items,
                                                            u64 *__restrict__ multi) {
  for (u64 r = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; r < n_reads;
       r += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 nk = kmer_ends[r] - (r ? kmer_ends[r - 1] : 0);
    const u64 n = nk ? (nk + item_kmers - 1) / item_kmers : 1;
    items[r] = n;
    multi[r] = n > 1 ? 1 : 0;
  }
}

// The batched form's color search: one wave per work item (a read, or up to
// item_kmers k-mers of a long one), over a device queue of (read, k-mer
// range) items, so a read of 100 kbp is spread over ~100 waves instead of one
// wave's chain of windows. A read of one item is thresholded by its wave. The
// items of a longer read add their counts into the read's slot (long_incl[r]
// - 1) of long_cnt, and their {found, not found, invalid} and a finished-item
// count into long_meta, and k_color_long_finish thresholds the slot after
// the launch (no fences in this kernel: a device-scope release per item would
// write back the XCD's L2).
template <bool LDS, int IN>
__global__ __launch_bounds__(256) void k_color_items(
    ColorView cv, const u64 *__restrict__ results, const u64 *__restrict__ kmer_ends, u64 n_reads,
    const u64 *__restrict__ item_ends, u32 item_kmers, const u64 *__restrict__ long_incl,
    KernelParams p, ReadOut o, u32 *__restrict__ global_cnt, u32 *__restrict__ long_cnt,
    u32 *__restrict__ long_meta, u64 *__restrict__ census) {
  extern __shared__ __align__(16) u32 lds[];
  uint4 *tab = wave_table(lds);
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const u32 waves = blockDim.x / kWave;
  const u32 nc = cv.num_colors;
  u32 *cnt = wave_counters<LDS>(lds, global_cnt, nc);
  if constexpr (LDS) {
    for (u32 c = lane; c < nc; c += kWave) cnt[c] = 0;
  }
  cnt_fence<LDS>();
  Census cs;
  cs.on = census != nullptr;
  u64 kmers = 0;
  const u64 n_items = item_ends[n_reads - 1];
  for (u64 it = static_cast<u64>(blockIdx.x) * waves + wave; it < n_items;
       it += static_cast<u64>(gridDim.x) * waves) {
    u64 lo = 0, hi = n_reads;  // the item's read: the first whose item end exceeds it
    while (lo < hi) {
      const u64 mid = (lo + hi) >> 1;
      if (item_ends[mid] > it) hi = mid;
      else lo = mid + 1;
    }
    const u64 r = lo;
    const u64 i0 = r ? item_ends[r - 1] : 0, ni = item_ends[r] - i0;
    const u64 rs = r ? kmer_ends[r - 1] : 0, re = kmer_ends[r];
    const u64 ks = min(re, rs + (it - i0) * item_kmers), ke = min(re, ks + item_kmers);
    u32 f = 0, nf = 0, inv = 0, touched = 0;
    read_add<IN>(cv, results, ks, ke, re, lane, cnt, tab, f, nf, inv, touched, cs);
    kmers += ke - ks;
    cnt_fence<LDS>();
    if (ni == 1) {
      finish_read<LDS>(cnt, nc, r, f, nf, inv, p, o, lane, touched);
    } else {
      const u64 slot = long_incl[r] - 1;
      u32 *lc = long_cnt + slot * nc;
      u32 *meta = long_meta + slot * 4;
      // the counter words this item wrote (all of them with global counters)
      const u32 tw = LDS ? wave_or(touched) : 0xFFFFFFFFu;
      for (u32 c = lane; c < nc; c += kWave) {
        if (LDS && !((tw >> (c / kWave)) & 1u)) continue;
        const u32 v = cnt_get<LDS>(cnt + c);
        if (v) {
          atomicAdd(lc + c, v);
          cnt[c] = 0;
        }
      }
      if (lane == 0) {
        atomicAdd(meta + 1, f);
        atomicAdd(meta + 2, nf);
        atomicAdd(meta + 3, inv);
      }
    }
    cnt_fence<LDS>();
  }
  census_flush(cs, kmers, lane, census);
}

// Reads of more than one work item, after k_color_items: one wave per such
// read thresholds its slot (every item's counts are in: the launch boundary
// orders them) and zeroes the slot for the next batch.
__global__ __launch_bounds__(256) void k_color_long_finish(u32 nc, u64 n_reads,
                                                            const u64 *__restrict__ multi,
                                                            const u64 *__restrict__ long_incl,
                                                            KernelParams p, ReadOut o,
                                                            u32 *__restrict__ long_cnt,
                                                            u32 *__restrict__ long_meta) {
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const u32 waves = blockDim.x / kWave;
  for (u64 r = static_cast<u64>(blockIdx.x) * waves + wave; r < n_reads;
       r += static_cast<u64>(gridDim.x) * waves) {
    if (!multi[r]) continue;
    const u64 slot = long_incl[r] - 1;
    u32 *meta = long_meta + slot * 4;
    const u32 f = meta[1], nf = meta[2], inv = meta[3];
    finish_read<false>(long_cnt + slot * nc, nc, r, f, nf, inv, p, o, lane, 0u);
    if (lane < 4) meta[lane] = 0;
  }
}

// Quirk 2 of the reference's colors stage: a read with no found index gets
// the colored-read slot of the next read that has one (colored_seq_id,
// src/IndexFileParser/IndexFileParser.cpp:65-74) and is printed with that
// slot's counts against its own threshold (ContinuousColorResultsPrinter.hpp
// :201-209). That only shows when its threshold is positive
// (--include-not-found / --include-invalid). colored_incl[r] = colored reads
// among 0..r; slot_read[q] = the read of colored slot q.
__global__ __launch_bounds__(256) void k_colored_slots(const u64 *__restrict__ colored_incl,
                                                        u64 n_reads, u64 *__restrict__ slot_read) {
  for (u64 r = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; r < n_reads;
       r += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 c = colored_incl[r], before = r ? colored_incl[r - 1] : 0;
    if (c != before) slot_read[c - 1] = r;
  }
}

// One wave per blank read with a positive threshold: the counts of the next
// colored read of the batch, thresholded with the blank read's minimum. A
// blank read with no colored read after it in the batch keeps its zero counts:
// the reference reads past its batch's counts there (slot = number of colored
// reads), so it has no defined output.
template <bool LDS, int IN>
__global__ __launch_bounds__(256) void k_color_quirk2(ColorView cv, const u64 *__restrict__ results,
                                                       const u64 *__restrict__ kmer_ends,
                                                       u64 n_reads, const u64 *__restrict__ blank_min,
                                                       const u64 *__restrict__ colored_incl,
                                                       const u64 *__restrict__ slot_read,
                                                       u64 *__restrict__ bitmap, u64 *__restrict__ nout,
                                                       u32 *__restrict__ global_cnt) {
  extern __shared__ __align__(16) u32 lds[];
  uint4 *tab = wave_table(lds);
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const u32 waves = blockDim.x / kWave;
  const u32 nc = cv.num_colors;
  u32 *cnt = wave_counters<LDS>(lds, global_cnt, nc);
  if constexpr (LDS) {
    for (u32 c = lane; c < nc; c += kWave) cnt[c] = 0;
  }
  cnt_fence<LDS>();
  const u32 words = (nc + 63) / 64;
  const u64 n_colored = colored_incl[n_reads - 1];
  for (u64 r = static_cast<u64>(blockIdx.x) * waves + wave; r < n_reads;
       r += static_cast<u64>(gridDim.x) * waves) {
    const u64 min_hits = blank_min[r];
    const u64 q = colored_incl[r];  // r found nothing: the colored reads before it
    if (min_hits == 0 || q >= n_colored) continue;
    const u64 rc = slot_read[q];
    const u64 ks = rc ? kmer_ends[rc - 1] : 0, ke = kmer_ends[rc];
    u32 f = 0, nf = 0, inv = 0;
    Census off;
    u32 touched = 0;
    read_add<IN>(cv, results, ks, ke, ke, lane, cnt, tab, f, nf, inv, touched, off);
    cnt_fence<LDS>();
    u64 total = 0;
    for (u32 w = 0; w < words; ++w) {
      const u32 c = w * 64 + lane;
      const u32 v = c < nc ? cnt_get<LDS>(cnt + c) : 0u;
      const bool hit = c < nc && v >= min_hits;
      const u64 m = __ballot(hit);
      if (lane == 0) bitmap[r * words + w] = m;
      total += __popcll(m);
      if (c < nc && v) cnt[c] = 0;
    }
    if (lane == 0) nout[r] = total;
    cnt_fence<LDS>();
  }
}

// Lanes per read of k_color_compact: the words of a read's hit bitmap side
// by side, a power of two up to a wave (several reads per wave when nc <= 2048).
inline u32 compact_lanes(u32 nc) {
  const u32 words = (nc + 63) / 64;
  u32 seg = 1;
  while (seg < words && seg < kWave) seg <<= 1;
  return seg;
}

// Expand per-read hit bitmaps into sorted u32 color ids at CSR offsets; ids
// past `capacity` are not written (offsets still give the full count). Each
// read takes `seg` lanes (compact_lanes), one bitmap word per lane and pass.
__global__ __launch_bounds__(256) void k_color_compact(const u64 *__restrict__ bitmap,
                                                        const u64 *__restrict__ incl, u64 n_reads,
                                                        u32 nc, u32 seg, u32 *__restrict__ ids,
                                                        u64 capacity, u64 *__restrict__ offsets) {
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, waves = blockDim.x / kWave;
  const u32 words = (nc + 63) / 64;
  const u32 per_wave = kWave / seg, sub = lane / seg, sl = lane % seg;
  for (u64 r0 = (static_cast<u64>(blockIdx.x) * waves + wave) * per_wave; r0 < n_reads;
       r0 += static_cast<u64>(gridDim.x) * waves * per_wave) {
    const u64 r = r0 + sub;
    const bool live = r < n_reads;
    u64 pos = live && r ? incl[r - 1] : 0;
    if (live && sl == 0) {
      offsets[r + 1] = incl[r];
      if (r == 0) offsets[0] = 0;
    }
    for (u32 w0 = 0; w0 < words; w0 += seg) {
      const u32 w = w0 + sl;
      u64 x = live && w < words ? bitmap[r * words + w] : 0;
      const u32 pc = static_cast<u32>(__popcll(x));
      u32 pre = pc;  // inclusive scan over the read's lanes
      for (u32 o = 1; o < seg; o <<= 1) {
        const u32 y = __shfl_up(pre, o, seg);
        if (sl >= o) pre += y;
      }
      u64 at = pos + pre - pc;
      while (x) {
        if (at < capacity) ids[at] = w * 64 + static_cast<u32>(__builtin_ctzll(x));
        ++at;
        x &= x - 1;
      }
      pos += __shfl(pre, sub * seg + seg - 1, kWave);
    }
  }
}

// Reference interface: one wave per colored read over its warps of 64
// warp-padded indexes (u64 max = padding), u64 counts per (read, color).
template <bool LDS>
__global__ __launch_bounds__(256) void k_color_warped(ColorView cv, const u64 *__restrict__ warped,
                                                       const u64 *__restrict__ intervals,
                                                       u64 n_colored, u64 *__restrict__ counts,
                                                       u32 *__restrict__ global_cnt) {
  extern __shared__ __align__(16) u32 lds[];
  uint4 *tab = wave_table(lds);
  const u32 lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, waves = blockDim.x / kWave;
  const u32 nc = cv.num_colors;
  u32 *cnt = wave_counters<LDS>(lds, global_cnt, nc);
  if constexpr (LDS) {
    for (u32 c = lane; c < nc; c += kWave) cnt[c] = 0;
  }
  cnt_fence<LDS>();
  for (u64 q = static_cast<u64>(blockIdx.x) * waves + wave; q < n_colored;
       q += static_cast<u64>(gridDim.x) * waves) {
    u32 f = 0, nf = 0, inv = 0;  // padding reads as not found
    const u64 wbeg = intervals[q] * kWave, wend = intervals[q + 1] * kWave;
    Census off;
    u32 touched = 0;
    read_add<kInKeys>(cv, warped, wbeg, wend, wend, lane, cnt, tab, f, nf, inv, touched, off);
    cnt_fence<LDS>();
    for (u32 c = lane; c < nc; c += kWave) {
      const u32 v = cnt_get<LDS>(cnt + c);
      counts[q * nc + c] = v;
      if (v) cnt[c] = 0;
    }
    cnt_fence<LDS>();
  }
}

inline u32 grid_for(u64 items, u32 per_block) {
  const u64 g = (items + per_block - 1) / per_block;
  return static_cast<u32>(std::max<u64>(1, std::min<u64>(g, kMaxGrid)));
}

// Workgroups of four waves; with global counters the grid is bounded so that
// each wave of it owns one slice of kGlobalCntBlocks * 4 slices.
constexpr u32 kColorWaves = 4;
constexpr u32 kGlobalCntBlocks = 1024;

inline bool lds_counters(u64 nc) { return nc <= kLdsColors; }

// Waves loop over their reads or items, so the grid is bounded: a few
// rounds of the chip's resident workgroups (fewer wave launches, counter
// clears and kernel prologues than one wave per read).
constexpr u32 kColorMaxBlocks = 16384;

inline u32 color_grid(u64 work_waves, u64 nc) {
  const u32 g = std::min<u32>(grid_for(work_waves, kColorWaves), kColorMaxBlocks);
  return lds_counters(nc) ? g : std::min<u32>(g, kGlobalCntBlocks);
}

// Each wave's table, then (nc <= kLdsColors) each wave's counters.
inline u32 color_lds_bytes(u64 nc) {
  return static_cast<u32>(kColorWaves * (kTabWords + (lds_counters(nc) ? nc : 0)) * 4);
}

template <typename F>
void with_counters(u64 nc, F f) {
  if (lds_counters(nc)) f.template operator()<true>();
  else f.template operator()<false>();
}

// The counter storage and the input kind as template arguments.
template <typename F>
void with_counters_input(u64 nc, int input, F f) {
  with_counters(nc, [&]<bool LDS>() {
    if (input == kInMasks) f.template operator()<LDS, kInMasks>();
    else if (input == kInRanks) f.template operator()<LDS, kInRanks>();
    else if (input == kInChainMasks) f.template operator()<LDS, kInChainMasks>();
    else if (input == kInChainRanks) f.template operator()<LDS, kInChainRanks>();
    else f.template operator()<LDS, kInKeys>();
  });
}

// The color table of an index: per node, the color entry of its key k-mer
// (the key walk of IndexSearcher.cuh:60-70 through the index's key
// destinations), as the fused color stage reads it. Ranks: the key's rank
// among the key k-mers (ColorSearcher.cuh:99-105); masks (at most
// kColorMaskMaxColors colors): the key's color set as a bit mask.
template <typename Idx, bool MASKS>
__global__ __launch_bounds__(256) void k_color_table(IndexView ix, ColorView cv, u64 n, void *out,
                                                      u64 *key_masks, u32 *marks_differ) {
  for (u64 v = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
       v += static_cast<u64>(gridDim.x) * blockDim.x) {
    // the lines' own marks rank the key k-mers the same way when they are the
    // color index's marks, bit for bit
    const ulonglong2 me = reinterpret_cast<const ulonglong2 *>(cv.key_marks)[v >> 6];
    const bool color_mark = ((me.y >> (v & 63)) & 1) != 0;
    const Idx b = block_of(static_cast<Idx>(v));
    const bool line_mark = entry_bit(*line_slot(ix, b, kSlotMarks), static_cast<u32>(v - b * kBlockNodes));
    if (line_mark != color_mark) *marks_differ = 1;
    u32 walked = 0;
    const u64 key = key_move<Idx>(ix, static_cast<Idx>(v), walked);
    const bool stuck = key == kWalkStuck;
    const u64 rank = stuck ? 0 : rank_marks(cv.key_marks, key);
    if constexpr (!MASKS) {
      static_cast<u32 *>(out)[v] = stuck ? kNoKeyRank : static_cast<u32>(rank);
    } else {
      u64 m = kWalkStuck;
      if (!stuck) {
        const uint4 e = cv.key_sets[rank];
        const u64 start = static_cast<u64>(e.x) | (static_cast<u64>(e.y) << 32);
        m = 0;
        if (e.w & 1u) {  // dense: bits [start, start + len), len <= num_colors < 64
          const u64 w = start >> 6;
          const u32 off = static_cast<u32>(start & 63);
          u64 x = cv.dense_arrays[w] >> off;
          if (off) x |= cv.dense_arrays[w + 1] << (64 - off);
          m = e.z >= 64 ? x : (x & ((1ULL << e.z) - 1));
        } else {
          for (u32 q = 0; q < e.z; ++q) {
            const u64 c = intv_get(cv.sparse_arrays, cv.sparse_width, start + q);
            if (c < cv.num_colors) m |= 1ULL << c;
          }
        }
        m &= (1ULL << cv.num_colors) - 1;
      }
      static_cast<u64 *>(out)[v] = m;
      if (color_mark) key_masks[rank] = m;  // a key k-mer (key == v): its mask by its rank
    }
  }
}

// key_sets[j] = set_table[color_set_idxs[j]], or an empty entry for an index
// past the sets. j runs to the number of keys inclusive: a node past the last
// key ranks there and reads the vector's zero padding word, as the
// reference's lookup does.
__global__ __launch_bounds__(256) void k_key_sets(const u64 *__restrict__ csi, u32 width,
                                                   const uint4 *__restrict__ table, u64 n_sets,
                                                   u64 n_keys, uint4 *__restrict__ out) {
  for (u64 j = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; j < n_keys;
       j += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 cs = intv_get(csi, width, j);
    out[j] = cs < n_sets ? table[cs] : make_uint4(0, 0, 0, 0);
  }
}

}  // namespace

void DeviceColors::build(const HostColors &c) {
  num_nodes = c.key_kmer_marks.bits;
  num_color_sets = c.num_color_sets;
  num_colors = c.num_colors;
  if (num_colors > 0xFFFFFFFFULL) throw std::runtime_error("too many colors");
  key_marks.build(c.key_kmer_marks.words.data(), c.key_kmer_marks.bits);
  csi_width = c.color_set_idxs.width;
  sparse_width = c.sparse_arrays.width;
  std::vector<u64> csi = c.color_set_idxs.words;
  csi.push_back(0);
  color_set_idxs = dev_upload(csi.data(), csi.size());
  std::vector<uint4> table(std::max<u64>(num_color_sets, 1));
  u64 dense_rank = 0;
  for (u64 cs = 0; cs < num_color_sets; ++cs) {
    u64 start, end;
    const bool dense = host_bit(c.is_dense_marks.words, cs);
    if (dense) {
      start = host_intv(c.dense_arrays_intervals, dense_rank);
      end = host_intv(c.dense_arrays_intervals, dense_rank + 1);
      ++dense_rank;
    } else {
      const u64 si = cs - dense_rank;
      start = host_intv(c.sparse_arrays_intervals, si);
      end = host_intv(c.sparse_arrays_intervals, si + 1);
    }
    if (end < start || end - start > 0xFFFFFFFFULL) throw std::runtime_error("malformed color set");
    uint4 e;
    e.x = static_cast<u32>(start);
    e.y = static_cast<u32>(start >> 32);
    e.z = static_cast<u32>(end - start);
    e.w = dense ? 1u : 0u;
    table[cs] = e;
  }
  set_table = dev_upload(table.data(), table.size());
  num_keys = c.color_set_idxs.size;
  SBWT_HIP(hipMalloc(&key_sets, (num_keys + 1) * sizeof(uint4)));
  k_key_sets<<<grid_for(num_keys + 1, 256), 256>>>(color_set_idxs, csi_width, set_table,
                                                    num_color_sets, num_keys + 1, key_sets);
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipDeviceSynchronize());
  std::vector<u64> dense = c.dense_arrays.words;
  dense.push_back(0);
  dense_arrays = dev_upload(dense.data(), dense.size());
  std::vector<u64> sparse = c.sparse_arrays.words;
  sparse.insert(sparse.end(), 2, 0);  // a chunk's 16-byte load may reach 12 bytes past its last id
  sparse_arrays = dev_upload(sparse.data(), sparse.size());
  device_bytes = key_marks.n_blocks * 16 + csi.size() * 8 + table.size() * 16 + (num_keys + 1) * 16 +
                 dense.size() * 8 + sparse.size() * 8;
}

void DeviceColors::release() {
  key_marks.release();
  void *bufs[] = {color_set_idxs, set_table, key_sets, dense_arrays, sparse_arrays};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  color_set_idxs = nullptr;
  set_table = nullptr;
  key_sets = nullptr;
  dense_arrays = nullptr;
  sparse_arrays = nullptr;
}

ColorView DeviceColors::view() const {
  ColorView v{};
  v.key_marks = key_marks.blocks;
  v.color_set_idxs = color_set_idxs;
  v.set_table = set_table;
  v.key_sets = key_sets;
  v.dense_arrays = dense_arrays;
  v.sparse_arrays = sparse_arrays;
  v.num_nodes = num_nodes;
  v.num_color_sets = num_color_sets;
  v.csi_width = csi_width;
  v.sparse_width = sparse_width;
  v.num_colors = static_cast<u32>(num_colors);
  return v;
}

ColorSearcher::ColorSearcher(const DeviceColors *colors, u64 max_indexes, u64 max_reads,
                             hipStream_t stream)
    : colors_(colors),
      stream_(stream),
      own_stream_(stream == nullptr),
      max_indexes_(max_indexes),
      max_reads_(max_reads) {
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  const u64 nc = colors_->num_colors;
  const u64 words = (nc + 63) / 64;
  SBWT_HIP(hipMalloc(&d_bitmap_, std::max<u64>(max_reads_ * words, 1) * 8));
  SBWT_HIP(hipMalloc(&d_nout_, std::max<u64>(max_reads_, 1) * 8 * 2));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(max_reads_) * 8));
  SBWT_HIP(hipMalloc(&d_quirk_, std::max<u64>(max_reads_, 1) * 8 * 4));
  if (!lds_counters(nc)) {  // one zeroed slice of counters per wave of the bounded grid
    const u64 bytes = u64{kGlobalCntBlocks} * kColorWaves * nc * 4;
    SBWT_HIP(hipMalloc(&d_global_cnt_, bytes));
    SBWT_HIP(hipMemset(d_global_cnt_, 0, bytes));
  }
  SBWT_HIP(hipMalloc(&d_census_, kCensusWords * 8));
  for (auto &e : ev_) SBWT_HIP(hipEventCreate(&e));
}

void ColorSearcher::last_census(u64 out[4]) {
  SBWT_HIP(hipMemcpyAsync(out, d_census_, kCensusWords * 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

void build_color_table(DeviceIndex &index, const DeviceColors &colors) {
  if (!index.has_marks) throw std::runtime_error("the index has no key-k-mer marks");
  if (colors.num_nodes != index.n) throw std::runtime_error("color index and SBWT differ in length");
  if (index.color_table) {
    SBWT_HIP(hipFree(index.color_table));
    index.color_table = nullptr;
  }
  if (index.key_color_masks) {
    SBWT_HIP(hipFree(index.key_color_masks));
    index.key_color_masks = nullptr;
  }
  index.color_ranks_in_lines = false;
  const bool masks = colors.num_colors <= kColorMaskMaxColors;
  if (!masks && colors.num_keys >= kNoKeyRank) throw std::runtime_error("too many key k-mers for a rank table");
  const u64 bytes = index.n * (masks ? 8 : 4) + 16;
  SBWT_HIP(hipMalloc(&index.color_table, bytes));
  u32 *d_differ = nullptr;
  SBWT_HIP(hipMalloc(&d_differ, sizeof(u32)));
  SBWT_HIP(hipMemset(d_differ, 0, sizeof(u32)));
  // masks: also each key k-mer's mask by its key rank (8 B per key)
  const u64 key_mask_bytes = masks ? (colors.num_keys + 1) * 8 : 0;
  if (masks) {
    SBWT_HIP(hipMalloc(&index.key_color_masks, key_mask_bytes));
    SBWT_HIP(hipMemset(index.key_color_masks, 0, key_mask_bytes));
  }
  const IndexView ix = index.view();
  const u32 grid = grid_for(index.n, 256);
  void *out = index.color_table;
  u64 *km = index.key_color_masks;
  if (index.wide) {
    if (masks) k_color_table<u64, true><<<grid, 256>>>(ix, colors.view(), index.n, out, km, d_differ);
    else k_color_table<u64, false><<<grid, 256>>>(ix, colors.view(), index.n, out, km, d_differ);
  } else {
    if (masks) k_color_table<u32, true><<<grid, 256>>>(ix, colors.view(), index.n, out, km, d_differ);
    else k_color_table<u32, false><<<grid, 256>>>(ix, colors.view(), index.n, out, km, d_differ);
  }
  SBWT_HIP(hipGetLastError());
  u32 differ = 1;
  SBWT_HIP(hipMemcpy(&differ, d_differ, sizeof(u32), hipMemcpyDeviceToHost));
  SBWT_HIP(hipFree(d_differ));
  index.color_table_kind = masks ? kColorTableMasks : kColorTableRanks;
  // the streaming search then hands a marked node's rank on in the result
  // after it (kRankShift), so the key pass need not read its node's entry
  index.color_ranks_in_lines = !index.wide && differ == 0 && colors.num_keys <= kRankMax;
  if (masks && !index.color_ranks_in_lines) {
    SBWT_HIP(hipFree(index.key_color_masks));
    index.key_color_masks = nullptr;
  } else {
    index.device_bytes += key_mask_bytes;
  }
  index.color_table_source = &colors;
  index.device_bytes += bytes;
}

void ColorSearcher::set_input(const DeviceIndex *index, bool chains) {
  chain_index_ = nullptr;
  if (!index) {
    input_ = kInKeys;
    return;
  }
  if (!index->color_table || index->color_table_source != colors_)
    throw std::runtime_error("the index's color table was not built from this color index");
  const bool masks = index->color_table_kind == kColorTableMasks;
  if (chains) {
    if (index->wide) throw std::runtime_error("key-chain input needs an index of fewer than 2^32 nodes");
    chain_index_ = index;
    input_ = masks ? kInChainMasks : kInChainRanks;
  } else {
    input_ = masks ? kInMasks : kInRanks;
  }
}

ColorView ColorSearcher::input_view() const {
  ColorView v = colors_->view();
  if (chain_index_) {
    v.chain_table = chain_index_->color_table;
    v.chain_key_masks = chain_index_->key_color_masks;
  }
  return v;
}

void ColorSearcher::set_work_items(u32 item_kmers) {
  if (item_kmers == 0) {
    item_kmers_ = 0;
    return;
  }
  if (item_kmers < kWave) throw std::runtime_error("work items hold at least 64 k-mers");
  const u64 nc = std::max<u64>(colors_->num_colors, 1);
  // a read of more than one item has more than item_kmers k-mers, so at most
  // max_indexes / item_kmers of them share a batch; their count slots (nc u32
  // each) are kept within a 32nd of device memory by lengthening the items
  size_t free_b = 0, total_b = 0;
  SBWT_HIP(hipMemGetInfo(&free_b, &total_b));
  const u64 budget = std::max<u64>(total_b / 32, 1 << 20);
  u64 ik = item_kmers;
  while ((max_indexes_ / ik + 1) * nc * 4 > budget && ik < (1ULL << 31)) ik *= 2;
  const u64 slots = max_indexes_ / ik + 1;
  if (slots > long_slots_ || d_items_ == nullptr) {
    for (void *b : {static_cast<void *>(d_items_), static_cast<void *>(d_long_cnt_),
                    static_cast<void *>(d_long_meta_)})
      if (b) SBWT_HIP(hipFree(b));
    SBWT_HIP(hipMalloc(&d_items_, std::max<u64>(max_reads_, 1) * 8 * 4));
    SBWT_HIP(hipMalloc(&d_long_cnt_, slots * nc * 4));
    SBWT_HIP(hipMalloc(&d_long_meta_, slots * 4 * 4));
    SBWT_HIP(hipMemset(d_long_cnt_, 0, slots * nc * 4));
    SBWT_HIP(hipMemset(d_long_meta_, 0, slots * 4 * 4));
    long_slots_ = slots;
  }
  item_kmers_ = static_cast<u32>(ik);
}

u32 ColorSearcher::autotune_items(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                                  const ColorParams &p, u64 *d_offsets, u32 *d_ids,
                                  u64 ids_capacity, float *best_ms) {
  static constexpr u32 kCandidates[] = {0, 256, 512, 1024, 2048, 4096};
  const bool timing = timing_;
  timing_ = true;
  float best = 0;
  const u32 before = item_kmers_;
  u32 best_items = before;
  bool any = false;
  try {
    // two runs per shape: the first warms the shape's buffers, the second is timed
    for (const u32 c : kCandidates) {
      set_work_items(c);
      pseudoalign_device(d_results, d_kmer_ends, n_reads, p, d_offsets, d_ids, ids_capacity,
                         nullptr);
      pseudoalign_device(d_results, d_kmer_ends, n_reads, p, d_offsets, d_ids, ids_capacity,
                         nullptr);
      const float ms = last_search_ms();
      if (!any || ms < best) {
        best = ms;
        best_items = c;
        any = true;
      }
    }
  } catch (...) {  // a failed batch leaves the searcher as it was
    timing_ = timing;
    set_work_items(before);
    throw;
  }
  set_work_items(best_items);
  timing_ = timing;
  if (best_ms) *best_ms = best;
  return item_kmers_;
}

ColorSearcher::~ColorSearcher() {
  for (auto &e : ev_) (void)hipEventDestroy(e);
  void *bufs[] = {d_bitmap_, d_nout_,      d_scan_tmp_, d_quirk_,    d_warped_,
                  d_intervals_, d_counts_, d_global_cnt_, d_items_, d_long_cnt_, d_long_meta_,
                  d_census_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

void ColorSearcher::search_warped_host(const u64 *warped, u64 n_idxs, const u64 *warp_intervals,
                                       u64 n_intervals, u64 *counts) {
  if (n_idxs == 0 || n_intervals < 2) return;
  if (n_idxs % kWave) throw std::runtime_error("warped indexes must be a multiple of 64");
  if (n_idxs > max_indexes_ || n_intervals - 1 > max_reads_)
    throw std::runtime_error("batch exceeds the color searcher's max indexes / reads");
  const u64 n_colored = n_intervals - 1;
  if (warp_intervals[n_colored] * kWave > n_idxs)
    throw std::runtime_error("warp intervals exceed the warped indexes");
  for (u64 q = 0; q < n_colored; ++q)
    if (warp_intervals[q + 1] < warp_intervals[q]) throw std::runtime_error("warp intervals decrease");
  const u64 nc = colors_->num_colors;
  if (d_warped_) SBWT_HIP(hipFree(d_warped_));
  if (d_intervals_) SBWT_HIP(hipFree(d_intervals_));
  if (d_counts_) SBWT_HIP(hipFree(d_counts_));
  SBWT_HIP(hipMalloc(&d_warped_, n_idxs * 8));
  SBWT_HIP(hipMalloc(&d_intervals_, n_intervals * 8));
  SBWT_HIP(hipMalloc(&d_counts_, std::max<u64>(n_colored * nc, 1) * 8));
  SBWT_HIP(hipMemcpyAsync(d_warped_, warped, n_idxs * 8, hipMemcpyHostToDevice, stream_));
  SBWT_HIP(hipMemcpyAsync(d_intervals_, warp_intervals, n_intervals * 8, hipMemcpyHostToDevice, stream_));
  SBWT_HIP(hipEventRecord(ev_[0], stream_));
  with_counters(nc, [&]<bool LDS>() {
    k_color_warped<LDS><<<color_grid(n_colored, nc), kColorWaves * kWave, color_lds_bytes(nc),
                          stream_>>>(colors_->view(), d_warped_, d_intervals_, n_colored, d_counts_,
                                     d_global_cnt_);
  });
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipEventRecord(ev_[1], stream_));
  recorded_ = true;
  SBWT_HIP(hipMemcpyAsync(counts, d_counts_, n_colored * nc * 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

u64 ColorSearcher::pseudoalign_device(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                                      const ColorParams &p, u64 *d_offsets, u32 *d_ids,
                                      u64 ids_capacity, u32 *d_stats) {
  if (n_reads == 0) return 0;
  if (n_reads > max_reads_) throw std::runtime_error("batch exceeds the color searcher's max reads");
  const u32 nc = static_cast<u32>(colors_->num_colors);
  launch_threshold(d_results, d_kmer_ends, n_reads, p, d_stats);
  u64 *d_incl = d_nout_ + max_reads_;
  scan_inclusive_u64(d_nout_, n_reads, d_incl, d_scan_tmp_, stream_);
  u64 total = 0;
  SBWT_HIP(hipMemcpyAsync(&total, d_incl + n_reads - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  if (total > ids_capacity)
    throw std::runtime_error("color id output needs " + std::to_string(total) + " entries, capacity " +
                             std::to_string(ids_capacity));
  const u32 seg = compact_lanes(nc), per_wave = kWave / seg;
  k_color_compact<<<grid_for((n_reads + per_wave - 1) / per_wave, 4), 4 * kWave, 0, stream_>>>(
      d_bitmap_, d_incl, n_reads, nc, seg, d_ids, ids_capacity, d_offsets);
  SBWT_HIP(hipGetLastError());
  return total;
}

void ColorSearcher::pseudoalign_device_async(const u64 *d_results, const u64 *d_kmer_ends,
                                             u64 n_reads, const ColorParams &p, u64 *d_offsets,
                                             u32 *d_ids, u64 ids_capacity, u32 *d_stats) {
  if (n_reads == 0) return;
  if (n_reads > max_reads_) throw std::runtime_error("batch exceeds the color searcher's max reads");
  const u32 nc = static_cast<u32>(colors_->num_colors);
  launch_threshold(d_results, d_kmer_ends, n_reads, p, d_stats);
  u64 *d_incl = d_nout_ + max_reads_;
  scan_inclusive_u64(d_nout_, n_reads, d_incl, d_scan_tmp_, stream_);
  const u32 seg = compact_lanes(nc), per_wave = kWave / seg;
  k_color_compact<<<grid_for((n_reads + per_wave - 1) / per_wave, 4), 4 * kWave, 0, stream_>>>(
      d_bitmap_, d_incl, n_reads, nc, seg, d_ids, ids_capacity, d_offsets);
  SBWT_HIP(hipGetLastError());
}

void ColorSearcher::launch_threshold(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                                     const ColorParams &p, u32 *d_stats) {
  const u64 nc = colors_->num_colors;
  KernelParams kp{p.threshold, p.include_not_found, p.include_invalid, p.first_read_quirk};
  // quirk 2 needs a positive threshold on a read that found nothing, which
  // only the include flags give
  const bool quirk2 = p.include_not_found || p.include_invalid;
  u64 *blank_min = d_quirk_, *colored = d_quirk_ + max_reads_;
  u64 *colored_incl = d_quirk_ + 2 * max_reads_, *slot_read = d_quirk_ + 3 * max_reads_;
  const ReadOut out{d_bitmap_, d_nout_, d_stats, quirk2 ? blank_min : nullptr,
                    quirk2 ? colored : nullptr};
  u64 *census = nullptr;
  if (census_on_) {  // measurement only: never on in a captured graph
    census = d_census_;
    SBWT_HIP(hipMemsetAsync(census, 0, kCensusWords * 8, stream_));
  }
  if (timing_) SBWT_HIP(hipEventRecord(ev_[0], stream_));
  if (item_kmers_) {
    // the batched form: a device queue of (read, k-mer range) items
    u64 *items = d_items_, *item_ends = d_items_ + max_reads_;
    u64 *multi = d_items_ + 2 * max_reads_, *long_incl = d_items_ + 3 * max_reads_;
    k_color_item_counts<<<grid_for(n_reads, 256), 256, 0, stream_>>>(d_kmer_ends, n_reads,
                                                                      item_kmers_, items, multi);
    scan_inclusive_u64(items, n_reads, item_ends, d_scan_tmp_, stream_);
    scan_inclusive_u64(multi, n_reads, long_incl, d_scan_tmp_, stream_);
    // items never outnumber reads plus k-mers / item_kmers; surplus waves leave
    const u64 bound = n_reads + max_indexes_ / item_kmers_ + 1;
    with_counters_input(nc, input_, [&]<bool LDS, int IN>() {
      k_color_items<LDS, IN><<<color_grid(bound, nc), kColorWaves * kWave, color_lds_bytes(nc),
                               stream_>>>(input_view(), d_results, d_kmer_ends, n_reads, item_ends,
                                      item_kmers_, long_incl, kp, out, d_global_cnt_, d_long_cnt_,
                                      d_long_meta_, census);
    });
    SBWT_HIP(hipGetLastError());
    k_color_long_finish<<<grid_for(n_reads, kColorWaves), kColorWaves * kWave, 0, stream_>>>(
        static_cast<u32>(nc), n_reads, multi, long_incl, kp, out, d_long_cnt_, d_long_meta_);
  } else {
    with_counters_input(nc, input_, [&]<bool LDS, int IN>() {
      k_color_reads<false, LDS, IN><<<color_grid(n_reads, nc), kColorWaves * kWave,
                                      color_lds_bytes(nc), stream_>>>(
          input_view(), d_results, d_kmer_ends, n_reads, kp, out, nullptr, d_global_cnt_, census);
    });
  }
  SBWT_HIP(hipGetLastError());
  if (quirk2) {
    scan_inclusive_u64(colored, n_reads, colored_incl, d_scan_tmp_, stream_);
    k_colored_slots<<<grid_for(n_reads, 256), 256, 0, stream_>>>(colored_incl, n_reads, slot_read);
    with_counters_input(nc, input_, [&]<bool LDS, int IN>() {
      k_color_quirk2<LDS, IN><<<color_grid(n_reads, nc), kColorWaves * kWave, color_lds_bytes(nc),
                                stream_>>>(input_view(), d_results, d_kmer_ends, n_reads, blank_min,
                                       colored_incl, slot_read, d_bitmap_, d_nout_, d_global_cnt_);
    });
    SBWT_HIP(hipGetLastError());
  }
  if (timing_) {
    SBWT_HIP(hipEventRecord(ev_[1], stream_));
    recorded_ = true;
  }
}

void ColorSearcher::counts_device(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                                  u32 *d_counts) {
  if (n_reads == 0) return;
  const u64 nc = colors_->num_colors;
  KernelParams kp{1.0, 0, 0, 0};
  const ReadOut out{nullptr, nullptr, nullptr, nullptr, nullptr};
  SBWT_HIP(hipEventRecord(ev_[0], stream_));
  with_counters_input(nc, input_, [&]<bool LDS, int IN>() {
    k_color_reads<true, LDS, IN><<<color_grid(n_reads, nc), kColorWaves * kWave,
                                   color_lds_bytes(nc), stream_>>>(input_view(), d_results, d_kmer_ends, n_reads, kp, out,
                                          d_counts, d_global_cnt_, nullptr);
  });
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipEventRecord(ev_[1], stream_));
  recorded_ = true;
}

float ColorSearcher::last_search_ms() {
  float ms = 0;
  if (!recorded_) return ms;
  SBWT_HIP(hipEventSynchronize(ev_[1]));
  SBWT_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  return ms;
}

}  // namespace sbwt_amd
