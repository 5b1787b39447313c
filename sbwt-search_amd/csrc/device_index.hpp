// Device-resident SBWT index: node lines (ACGT rank entries, suffix-group
// starts and key-k-mer marks per 96 nodes) and the presearch table. Replaces the reference's GpuSbwtContainer
// (src/SbwtContainer/GpuSbwtContainer.cpp:16-52) and Presearcher
// (src/Presearcher/Presearcher.cpp:18-30) with an MI355X layout.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace sbwt_amd {

// Mark entries of a single bit vector: per 64 bits one 16-byte entry
// {u64 ones before it, the 64 bits}. rank over [0, i) = one 16-byte load.
struct DeviceMarks {
  u64 *blocks = nullptr;  // [n_blocks][2]
  u64 n_bits = 0, n_blocks = 0, total = 0;
  void build(const u64 *words, u64 n_bits);
  void release();
};

struct DeviceIndex {
  int device = 0;
  u64 n = 0;         // nodes
  u32 k = 0;
  u32 kp = 0;        // presearch prefix length
  bool wide = false; // node indexes need 64 bits (n >= 2^32)
  u64 c_map[5] = {0, 0, 0, 0, 0};
  uint4 *blocks = nullptr;  // [n_blocks][8] slots: one 128-byte line per 96 nodes
  u64 n_blocks = 0;
  u64 *hyper_base = nullptr;
  u64 n_hyper = 0;
  u32 hyper_shift = kHyperShift;
  void *presearch = nullptr;
  u64 presearch_bytes = 0;
  bool has_marks = false;   // key-k-mer marks are in the lines
  void *key_dest = nullptr; // per node: its key-k-mer move destination (u32, or u64 if wide)
  uint4 *kmer_hash = nullptr;  // k-mer code -> node table (narrow indexes), or null
  u32 kmer_hash_bits = 0;       // group index bits
  u32 kmer_hash_node_bits = 0;  // entry bits of node + 1
  // per node: its key's color entry (masks or key ranks), built from a color
  // index by build_color_table (color_search.hip); searchers use it only when
  // asked to output color-table entries
  void *color_table = nullptr;
  u32 color_table_kind = kColorTableNone;
  const void *color_table_source = nullptr;  // the DeviceColors it was built from
  // the color table was built over marks equal to the lines' own (checked
  // when it is built): a marked node's line rank is its key rank, its entry
  // with key ranks, the index of its mask in key_color_masks with masks
  bool color_ranks_in_lines = false;
  u64 *key_color_masks = nullptr;  // masks color table: [key rank] = the key's mask
  bool has_groups = false;  // suffix-group starts are in the lines
  // starts were given but some non-start row has an edge: not an SBWT, so
  // every search runs per k-mer (the streaming search needs the SBWT property)
  bool groups_rejected = false;
  u64 device_bytes = 0;

  // acgt: four host bit vectors of ceil(n/64) words (bit i of word w = node
  // 64w + i, as in the .sbwt file). key_marks and sgs_words (the suffix-group
  // starts, which let the streaming search step from any node) may be null.
  void build(const u64 *const acgt[4], u64 n, u32 k, u32 kp, const u64 *key_marks_words,
             u64 key_marks_bits, const u64 *sgs_words);
  void release();
  IndexView view() const;
  // The k-mer table that lets a fresh search read about one bucket pair
  // instead of the presearch entry and k - kp steps; built on request (about
  // 0.1 s for 10^8 nodes), used by every later search. False when the
  // index cannot have one (64-bit nodes, k > 32, not enough memory).
  bool build_kmer_hash();
};

// Presearch prefix length used when the caller does not choose one. The
// reference fixes 12 letters (src/Global/GlobalDefinitions.h.in:17); here the
// prefix grows with the index to about log4(n) + 2 letters (between 12 and 16,
// at most k), so that a prefix interval holds a node or none and most searches
// start with few SBWT steps left, as long as the table takes at most an eighth
// of device memory (8 bytes per prefix, 16 for 64-bit indexes).
u32 default_presearch_letters(u32 k, u64 n, bool wide, u64 device_bytes);

}  // namespace sbwt_amd
