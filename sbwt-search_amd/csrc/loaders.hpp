// Readers of the on-disk index formats the reference consumes:
//  .sbwt/.tdbg "plain-matrix" v0.1 (src/SbwtBuilder/SbwtBuilder.cpp:58-145)
//  .tcolors "sdsl-hybrid-v4" (src/ColorIndexBuilder/ColorIndexBuilder.cpp:23-86)
// sdsl-lite c32874c serialisation: bit_vector = u64 bits + words;
// int_vector<> = u64 bits + u8 width + words; rank_support_v5 = int_vector<64>.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sbwt_amd {

struct HostSbwt {
  uint64_t num_bits = 0;
  uint64_t k = 0;
  std::vector<uint64_t> acgt[4];
  std::vector<uint64_t> suffix_group_starts;  // bit i set: node i starts a suffix group
};

struct HostIntVector {
  std::vector<uint64_t> words;
  uint64_t size = 0;  // elements
  uint32_t width = 0;
};

struct HostBitVector {
  std::vector<uint64_t> words;
  uint64_t bits = 0;
};

struct HostColors {
  HostBitVector dense_arrays;
  HostIntVector dense_arrays_intervals;
  HostIntVector sparse_arrays;
  HostIntVector sparse_arrays_intervals;
  HostBitVector is_dense_marks;
  HostBitVector key_kmer_marks;
  HostIntVector color_set_idxs;
  uint64_t num_color_sets = 0;
  uint64_t num_colors = 0;
};

HostSbwt load_sbwt(const std::string &path);

// Whether a suffix-group-starts vector fits the SBWT it came with: in an SBWT
// only the first node of each suffix group (the nodes sharing their last k - 1
// letters) carries the group's outgoing edges, so every node that is not
// marked as a group start has an empty row in all four letter vectors. This is
// what the reference's interval step C[c] + rank_c relies on to count each
// edge once (src/IndexSearcher/IndexSearcher.cuh:48-55), and what a search
// stepping from a single node through its group start needs. Bits past
// num_bits are ignored; node 0 (the root) always starts a group.
bool suffix_groups_consistent(const uint64_t *const acgt[4], const uint64_t *suffix_group_starts,
                              uint64_t num_bits);
HostBitVector load_key_kmer_marks(const std::string &tcolors_path);
HostColors load_colors(const std::string &tcolors_path);

}  // namespace sbwt_amd
