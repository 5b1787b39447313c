// Shared definitions of the MI355X (gfx950) search library: error handling,
// device-side index views and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace sbwt_amd {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr u64 kNotFound = ~0ULL;        // reference "-1" (IndexSearcher.cuh:57)
constexpr u64 kInvalid = ~0ULL - 1;     // reference "-2" (BinaryContinuousIndexResultsPrinter)
constexpr u64 kNewline = ~0ULL - 2;     // end of a read in the binary index format (same file)
constexpr u64 kWalkStuck = ~0ULL - 3;   // key walk reached a row with no edge and no mark
// Key-destination table entries that are not node indexes: the walk from the
// node is longer than the table's build cap (walk at query time), or stuck.
constexpr u32 kKeyLong32 = 0xFFFFFFFFu, kKeyStuck32 = 0xFFFFFFFEu;
constexpr u64 kKeyLong64 = ~0ULL;
constexpr u32 kKeyTableSteps = 4096;  // walk cap of the table build

// HIP failures become std::runtime_error("GPUassert: <msg> at <file>:<line>"),
// the message shape of the reference's GPU_CHECK (src/Tools/GpuUtils.h:20-29).
[[noreturn]] void throw_hip(hipError_t e, const char *file, int line);
#define SBWT_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) ::sbwt_amd::throw_hip(e_, __FILE__, __LINE__); \
  } while (0)

// ---------------------------------------------------------------------------
// Device layout of the SBWT ("node lines").
// Line b covers nodes [96 b, 96 b + 96) and is one 128-byte line (the L2 line
// of the MI355X, so every query on a node costs one line) of eight 16-byte
// slots:
//   slot c (0..3)  letter c: {count, bits[0..31], bits[32..63], bits[64..95]}
//                  count = 1s of letter c before the line
//   slot 4         suffix-group starts: {delta, bits[0..95]}, delta = distance
//                  from the line's first node back to the start of its suffix
//                  group (kFarGroup if unknown or too far)
//   slot 5         key-k-mer marks: {count, bits[0..95]}
//   slots 6, 7     key-walk codes of nodes 0..47 and 48..95 (indexes with key
//                  marks; zero otherwise): one of six values per node, three
//                  nodes per byte (c0 + 6 c1 + 36 c2): 0..3 the node is not
//                  marked and its key walk takes its first edge, that letter;
//                  4 the node is marked; 5 neither (no edge: the walk is stuck)
// Counts are relative to the line's hyperblock (2^25 lines) whose absolute
// bases sit in a small table; indexes below 2^32 nodes have a single
// hyperblock with base 0, so a rank is one 16-byte load and a popcount.
// The color index keeps its key marks as 16-byte "mark entries": {u64 count,
// 64 bits}.
// ---------------------------------------------------------------------------
constexpr u32 kBlockNodes = 96;
constexpr u32 kLineSlots = 8;
constexpr u32 kSlotGroups = 4;
constexpr u32 kSlotMarks = 5;
constexpr u32 kSlotWalk = 6;            // slots 6 and 7
constexpr u32 kWalkNodesPerSlot = 48;
constexpr u32 kWalkMarked = 4, kWalkNoEdge = 5;
// What the streaming search leaves in bits 61-62 of a found k-mer's node for
// the key pass, about the k-mer before it in the read (found, same work item):
// nothing known, same key as this k-mer (its walk's first step is this
// k-mer's node), marked (its own key), or stuck walk.
constexpr u32 kPredShift = 61;
constexpr u32 kPredLookup = 0, kPredSameKey = 1, kPredMarked = 2, kPredStuck = 3;
// With a color table built over the lines' own marks (IndexView::line_ranks,
// narrow indexes), bits 32-60 of such a result hold 1 + the rank of the
// marked node before it among the marked nodes, read from the same line (0
// when that node is not marked or the rank does not fit): its color-table
// entry with key ranks, the index of its mask in key_masks with masks.
constexpr u32 kRankShift = 32;
constexpr u32 kRankMax = (1u << 29) - 2;  // largest rank the field holds (as rank + 1)
constexpr u32 kHyperShift = 25;  // lines per hyperblock = 2^25 (3.2 G nodes)
constexpr u32 kMarkBits = 64;  // bits per mark entry of the color index
constexpr u32 kFarGroup = 0xFFFFFFFFu;

struct IndexView {
  const uint4 *blocks;        // [n_blocks][8] slots
  const u64 *hyper_base;      // [n_hyper][8] absolute counts at hyperblock start (slots 0-3, 5)
  u32 hyper_shift;            // lines per hyperblock = 2^hyper_shift (kHyperShift)
  const void *presearch;      // [4^kp] of {left,right} (uint2 or ulonglong2)
  const void *key_dest;       // [n] key-k-mer move destination (u32 or u64), or null
  const uint4 *kmer_hash;     // [2^kmer_hash_bits groups][2] (four 8-byte entries each), or null
  u32 kmer_hash_bits;         // G: group index bits
  u32 kmer_hash_node_bits;    // N: entry bits of node + 1
  u64 c_map[5];
  u64 num_nodes;
  u32 k;
  u32 kp;
  u32 has_groups;             // slot 4 holds the suffix-group starts
  u32 has_marks;              // slot 5 holds the key-k-mer marks
  // With the key-k-mer move, results may be the color-table entry of each
  // k-mer's key instead of the key's node: per node, its key's color mask
  // (kColorTableMasks) or its key's rank among the key k-mers
  // (kColorTableRanks); null for key nodes.
  const void *color_table;
  u32 color_table_kind;
  u32 line_ranks;             // color table keyed like the lines' mark ranks (kRankShift)
  const u64 *key_masks;       // masks color table: the mask of each key k-mer, by key rank
};

constexpr u32 kColorTableNone = 0, kColorTableMasks = 1, kColorTableRanks = 2;
// What a searcher with the key move writes for a found k-mer: its key node,
// its key's color-table entry, or (key chains, narrow indexes) its node with
// what the search learnt about the k-mer before it (kPredShift, kRankShift),
// for a color searcher to resolve.
constexpr int kColorOutputNone = 0, kColorOutputEntries = 1, kColorOutputChains = 2;
constexpr u32 kColorMaskMaxColors = 61;  // masks leave bits 61-63 to the sentinels
constexpr u32 kNoKeyRank = 0xFFFFFFFFu;  // rank table entry of a node whose key walk is stuck

// k-mer table: every node whose label is a whole k-mer (k <= 32, fewer than
// 2^32 nodes), in groups of four 8-byte entries (32 bytes, a quarter of a
// 128-byte line). A k-mer's 2k-bit code times an odd constant modulo 4^k (a
// bijection) is its mix; the top G bits of the mix pick its home group, and
// its entry holds the remaining 2k - G bits (the remainder), the distance d
// (0-3) from the home group to the group holding it, and node + 1 in the top
// N bits (2^N > nodes, so a used entry is never 0):
//   entry = (node + 1) << (64 - N) | remainder << 2 | d,  2 + (2k - G) <= 64 - N.
// An insert takes the first free entry of groups home ... home + 3, in entry
// order (a k-mer that finds none stays out); a lookup reads those groups in
// turn, one 32-byte group each, and stops at the k-mer's entry, at an empty
// entry (the k-mer is absent: it would have been put there or before) or
// after kHashGroups groups (the full search answers).
constexpr u32 kHashGroups = 4;
__host__ __device__ __forceinline__ u64 kmer_hash_mix(u64 code, u32 k) {
  const u64 m = code * 0x9E3779B97F4A7C15ULL;
  return k >= 32 ? m : m & ((1ULL << (2 * k)) - 1);
}
// The low 64 - N bits of the k-mer's entry in the group d after its home
// (rem_bits = 2k - G).
__host__ __device__ __forceinline__ u64 kmer_hash_tag(u64 mix, u32 rem_bits, u32 d) {
  return ((mix & ((1ULL << rem_bits) - 1)) << 2) | d;
}

// Words of the search census (sbwt_searcher_last_stats).
constexpr u32 kStatsWords = 16;

// Threads per workgroup of the search kernels: 4 wave64s.
constexpr u32 kSearchThreads = 256;

}  // namespace sbwt_amd
