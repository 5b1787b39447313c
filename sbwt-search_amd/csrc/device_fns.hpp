// Device functions over the MI355X index layout: rank on node lines and mark
// blocks, the SBWT step, bit reads and sdsl int_vector reads.
#pragma once

#include "common.hpp"

namespace sbwt_amd {

// Number of 1s among the first `o` (0..95) bits of a rank-block entry.
__device__ __forceinline__ u32 entry_popc(const uint4 e, u32 o) {
  const u64 lo = (static_cast<u64>(e.z) << 32) | e.y;
  const u64 mlo = o >= 64 ? ~0ULL : ((1ULL << o) - 1);
  const u32 mhi = o > 64 ? ((1u << (o - 64)) - 1) : 0u;
  return static_cast<u32>(__popcll(lo & mlo)) + static_cast<u32>(__popc(e.w & mhi));
}

__device__ __forceinline__ bool entry_bit(const uint4 e, u32 o) {
  const u32 w = o < 32 ? e.y : (o < 64 ? e.z : e.w);
  return (w >> (o & 31)) & 1u;
}

template <typename Idx>
__device__ __forceinline__ Idx block_of(Idx i) {
  return i / static_cast<Idx>(kBlockNodes);
}

// Absolute count before line b of slot `slot` (letters 0..3, marks 5):
// narrow indexes store absolute counts; wide ones add the hyperblock base.
template <typename Idx>
__device__ __forceinline__ Idx entry_base(const IndexView &ix, uint4 e, Idx b, u32 slot) {
  if constexpr (sizeof(Idx) == 4) {
    return e.x;
  } else {
    return ix.hyper_base[(b >> ix.hyper_shift) * kLineSlots + slot] + e.x;
  }
}

template <typename Idx>
__device__ __forceinline__ Idx cmap_of(const IndexView &ix, u32 c) {
  const u64 v = c == 0 ? ix.c_map[0] : c == 1 ? ix.c_map[1] : c == 2 ? ix.c_map[2] : ix.c_map[3];
  return static_cast<Idx>(v);
}

__device__ __forceinline__ const uint4 *line_slot(const IndexView &ix, u64 b, u32 slot) {
  return ix.blocks + b * kLineSlots + slot;
}

// rank_c over [0, i): src/UtilityKernels/Rank.cuh:26-65 restated on the
// node-line layout (one 16-byte load).
template <typename Idx>
__device__ __forceinline__ Idx rank_acgt(const IndexView &ix, u32 c, Idx i) {
  const Idx b = block_of(i);
  const u32 o = static_cast<u32>(i - b * kBlockNodes);
  const uint4 e = *line_slot(ix, b, c);
  return entry_base(ix, e, b, c) + entry_popc(e, o);
}

// One SBWT step with letter c on the interval [l, r]
// (src/IndexSearcher/IndexSearcher.cuh:48-55). When l and r+1 fall in the
// same line the entry is loaded once.
template <typename Idx>
__device__ __forceinline__ void sbwt_step(const IndexView &ix, u32 c, Idx &l, Idx &r) {
  const Idx cm = cmap_of<Idx>(ix, c);
  const Idx r1 = r + 1;
  const Idx bl = block_of(l), br = block_of(r1);
  const uint4 el = *line_slot(ix, bl, c);
  uint4 er = el;
  if (br != bl) er = *line_slot(ix, br, c);
  l = cm + entry_base(ix, el, bl, c) + entry_popc(el, static_cast<u32>(l - bl * kBlockNodes));
  r = cm + entry_base(ix, er, br, c) + entry_popc(er, static_cast<u32>(r1 - br * kBlockNodes)) - 1;
}

// rank over [0, i) on mark entries {u64 count, u64 bits} (one 16-byte load).
__device__ __forceinline__ u64 rank_marks(const u64 *mb, u64 i) {
  const ulonglong2 e = reinterpret_cast<const ulonglong2 *>(mb)[i >> 6];
  return e.x + __popcll(e.y & ((1ULL << (i & 63)) - 1));
}

// Key-k-mer move (src/IndexSearcher/IndexSearcher.cuh:60-70): while node l
// is not a key k-mer, follow its first outgoing edge in A,C,G,T order. A row
// with no edge ends the walk with kWalkStuck (the reference loops forever).
// Each step reads one node line: the marks slot and the four letter slots.
template <typename Idx>
__device__ __forceinline__ u64 key_walk(const IndexView &ix, Idx l, u64 max_steps, u32 &walked) {
  for (u64 s = 0; s <= max_steps; ++s, ++walked) {
    const Idx b = block_of(l);
    const u32 o = static_cast<u32>(l - b * kBlockNodes);
    const uint4 *line = line_slot(ix, b, 0);
    const uint4 m = line[kSlotMarks];
    const uint4 e0 = line[0], e1 = line[1], e2 = line[2], e3 = line[3];
    if (entry_bit(m, o)) return static_cast<u64>(l);
    u32 c = 4;
    uint4 e = e0;
    if (entry_bit(e3, o)) { c = 3; e = e3; }
    if (entry_bit(e2, o)) { c = 2; e = e2; }
    if (entry_bit(e1, o)) { c = 1; e = e1; }
    if (entry_bit(e0, o)) { c = 0; e = e0; }
    if (c == 4) return kWalkStuck;
    l = cmap_of<Idx>(ix, c) + entry_base(ix, e, b, c) + entry_popc(e, o);
  }
  return kWalkStuck;
}

// The key walk of node l, at most max_steps steps; kKeyLong64 when the cap
// ends it (the table build records such nodes for a walk at query time).
template <typename Idx>
__device__ __forceinline__ u64 key_walk_capped(const IndexView &ix, Idx l, u32 max_steps) {
  for (u32 s = 0; s < max_steps; ++s) {
    const Idx b = block_of(l);
    const u32 o = static_cast<u32>(l - b * kBlockNodes);
    const uint4 *line = line_slot(ix, b, 0);
    const uint4 m = line[kSlotMarks];
    const uint4 e0 = line[0], e1 = line[1], e2 = line[2], e3 = line[3];
    if (entry_bit(m, o)) return static_cast<u64>(l);
    u32 c = 4;
    uint4 e = e0;
    if (entry_bit(e3, o)) { c = 3; e = e3; }
    if (entry_bit(e2, o)) { c = 2; e = e2; }
    if (entry_bit(e1, o)) { c = 1; e = e1; }
    if (entry_bit(e0, o)) { c = 0; e = e0; }
    if (c == 4) return kWalkStuck;
    l = cmap_of<Idx>(ix, c) + entry_base(ix, e, b, c) + entry_popc(e, o);
  }
  return kKeyLong64;
}

// Decoded key-destination table entry of node l: its key node, kWalkStuck, or
// kKeyLong64 (walk at query time).
template <typename Idx>
__device__ __forceinline__ u64 key_dest_entry(const IndexView &ix, Idx l) {
  if constexpr (sizeof(Idx) == 4) {
    const u32 d = static_cast<const u32 *>(ix.key_dest)[l];
    return d == kKeyLong32 ? kKeyLong64 : (d == kKeyStuck32 ? kWalkStuck : static_cast<u64>(d));
  } else {
    return static_cast<const u64 *>(ix.key_dest)[l];
  }
}

// The key-k-mer move of a found node: one table read, or the walk when the
// index has no table or the node's walk is longer than its cap.
template <typename Idx>
__device__ __forceinline__ u64 key_move(const IndexView &ix, Idx l, u32 &walked) {
  if (ix.key_dest) {
    ++walked;
    const u64 d = key_dest_entry(ix, l);
    if (d != kKeyLong64) return d;
  }
  return key_walk<Idx>(ix, l, ix.num_nodes, walked);
}

// The color-table entry of node l's key (IndexView::color_table): its key's
// color mask or key rank; kWalkStuck when the key walk has no end.
template <typename Idx>
__device__ __forceinline__ u64 color_table_entry(const IndexView &ix, Idx l) {
  if (ix.color_table_kind == kColorTableMasks) return static_cast<const u64 *>(ix.color_table)[l];
  const u32 r = static_cast<const u32 *>(ix.color_table)[l];
  return r == kNoKeyRank ? kWalkStuck : static_cast<u64>(r);
}

// The key move of a found node l, or its key's color-table entry when the
// view carries a color table (one line either way for most nodes).
template <typename Idx>
__device__ __forceinline__ u64 key_result(const IndexView &ix, Idx l, u32 &walked) {
  if (ix.color_table) {
    ++walked;
    return color_table_entry<Idx>(ix, l);
  }
  return key_move<Idx>(ix, l, walked);
}

// sdsl int_vector<> element i of width w (LSB-first, may straddle two words):
// src/UtilityKernels/VariableLengthIntegerIndex.cuh:23-36.
__device__ __forceinline__ u64 intv_get(const u64 *v, u32 width, u64 i) {
  const u64 bit = i * width, w = bit >> 6;
  const u32 off = static_cast<u32>(bit & 63);
  u64 x = v[w] >> off;
  if (off + width > 64) x |= v[w + 1] << (64 - off);
  return width == 64 ? x : (x & ((1ULL << width) - 1));
}

__device__ __forceinline__ bool plain_bit(const u64 *bv, u64 i) { return (bv[i >> 6] >> (i & 63)) & 1ULL; }

}  // namespace sbwt_amd
