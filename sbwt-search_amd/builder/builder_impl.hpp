// Internal to libsbwt_build.so: the SBWT and color-index structures that the
// host builder (builder.cpp) and the GPU builder (gpu_build.hip) both fill and
// that the writers in builder.cpp serialize.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace sbwt_build {

inline uint64_t low_mask(uint32_t digits) {  // low 2*digits bits set
  return digits >= 32 ? ~0ULL : ((1ULL << (2 * digits)) - 1);
}

inline int bits_for(uint64_t max_value) {  // sdsl bits::hi(x)+1, at least 1
  return max_value == 0 ? 1 : 64 - __builtin_clzll(max_value);
}

}  // namespace sbwt_build

// Nodes are the k-mers (real) plus the dummy nodes $^(k-j) p[0..j-1] needed
// for every k-mer with no predecessor, in colexicographic order ($ < A < C <
// G < T). A node is (primary, len): primary holds char i of the node string at
// bits [2i, 2i+1] with $ as 0, len = number of non-$ chars; colex order is
// (primary, len) ascending.
struct bld_sbwt {
  uint64_t n = 0, n_kmers = 0;
  uint32_t k = 0;
  bool add_rc = true;
  std::vector<uint64_t> kmers;      // sorted colex keys of the real k-mers (may be absent)
  std::vector<uint64_t> dummy_key;  // sorted dummies
  std::vector<uint8_t> dummy_len;
  std::vector<uint64_t> bits[5];    // A, C, G, T, suffix-group starts
  uint64_t c_map[5] = {0, 0, 0, 0, 0};

  // number of dummies ordered before real k-mer key x
  uint64_t dummies_before(uint64_t x) const {
    return std::upper_bound(dummy_key.begin(), dummy_key.end(), x) - dummy_key.begin();
  }
  uint64_t node_of_kmer_index(uint64_t i) const { return i + dummies_before(kmers[i]); }
  bool has_kmers() const { return kmers.size() == n_kmers; }
};

struct bld_colors {
  uint64_t n = 0, num_colors = 0;
  std::vector<uint64_t> key_marks;
  std::vector<uint64_t> set_offsets;  // set s = set_ids[set_offsets[s] .. set_offsets[s+1])
  std::vector<uint32_t> set_ids;
  std::vector<uint64_t> color_set_idxs;  // one per key node, in node order
  uint64_t n_dense = 0;
  std::vector<uint8_t> dense;
};

namespace sbwt_build {
// Dense flags of every set: a set is dense when its bitmap up to its largest
// id is smaller than its id list at the sparse width.
inline void choose_dense(bld_colors *C) {
  const uint64_t nsets = C->set_offsets.size() - 1;
  const int sparse_w = bits_for(C->num_colors - 1);
  C->dense.assign(nsets, 0);
  C->n_dense = 0;
  for (uint64_t i = 0; i < nsets; ++i) {
    const uint64_t sz = C->set_offsets[i + 1] - C->set_offsets[i];
    const uint64_t mx = sz ? C->set_ids[C->set_offsets[i + 1] - 1] : 0;
    if (sz && mx + 1 < sz * static_cast<uint64_t>(sparse_w)) {
      C->dense[i] = 1;
      C->n_dense++;
    }
  }
}
}  // namespace sbwt_build
