// On-disk index readers (see loaders.hpp).
#include "loaders.hpp"

#include <fstream>
#include <stdexcept>

namespace sbwt_amd {

namespace {

struct In {
  std::ifstream f;
  std::string path;
  explicit In(const std::string &p) : f(p, std::ios::in | std::ios::binary), path(p) {
    if (!f) throw std::runtime_error("The input file " + p + " cannot be opened");  // IOUtils.cpp:16-21
  }
  void raw(void *dst, uint64_t n) {
    f.read(static_cast<char *>(dst), static_cast<std::streamsize>(n));
    if (static_cast<uint64_t>(f.gcount()) != n) throw std::runtime_error("Unexpected end of file " + path);
  }
  uint64_t u64() {
    uint64_t v;
    raw(&v, 8);
    return v;
  }
  // u64 length + bytes: src/Tools/IOUtils.cpp:44-56
  std::string str() {
    const uint64_t n = u64();
    if (n > (1u << 20)) throw std::runtime_error("Malformed string in " + path);
    std::string s(n, '\0');
    raw(s.data(), n);
    return s;
  }
  void skip(uint64_t n) { f.seekg(static_cast<std::streamoff>(n), std::ios::cur); }
  void words(std::vector<uint64_t> &w, uint64_t bits) {
    const uint64_t n = (bits + 63) / 64;
    w.assign(n + 1, 0);  // one zero word of padding (GpuColorIndexContainer.cpp:4-9)
    raw(w.data(), n * 8);
  }
  void bit_vector(HostBitVector &b) {
    b.bits = u64();
    words(b.words, b.bits);
  }
  void int_vector(HostIntVector &v) {
    const uint64_t bits = u64();
    uint8_t w;
    raw(&w, 1);
    if (w == 0 || w > 64) throw std::runtime_error("Malformed int_vector width in " + path);
    v.width = w;
    v.size = bits / w;
    words(v.words, bits);
  }
  void skip_bit_vector() {
    const uint64_t bits = u64();
    skip((bits + 63) / 64 * 8);
  }
  void skip_int_vector() {
    const uint64_t bits = u64();
    skip(1 + (bits + 63) / 64 * 8);
  }
};

void check_colors_magic(In &in) {
  if (in.str() != "sdsl-hybrid-v4")
    throw std::runtime_error("The colors file has an incorrect format. Expected 'sdsl-hybrid-v4'");
}

}  // namespace

HostSbwt load_sbwt(const std::string &path) {
  In in(path);
  HostSbwt s;
  const std::string variant = in.str();
  if (variant != "v0.1") {  // SbwtBuilder.cpp:61-68
    if (variant != "plain-matrix") throw std::runtime_error("Error input is not a plain-matrix SBWT");
    if (in.str() != "v0.1") throw std::runtime_error("Error: wrong SBWT version");
  }
  for (int c = 0; c < 4; ++c) {
    const uint64_t bits = in.u64();
    if (c == 0) s.num_bits = bits;
    else if (bits != s.num_bits) throw std::runtime_error("Malformed SBWT: bit vectors differ in length");
    in.words(s.acgt[c], bits);
  }
  for (int i = 0; i < 4; ++i) in.skip_bit_vector();  // rank structures (SbwtBuilder.cpp:97)
  {  // suffix group starts
    const uint64_t bits = in.u64();
    if (bits != s.num_bits) throw std::runtime_error("Malformed SBWT: suffix group starts length");
    in.words(s.suffix_group_starts, bits);
  }
  in.skip(in.u64());                                  // C map
  in.skip(in.u64());                                  // kmer_prefix_calc
  in.skip(24);                                        // precalc_k, n_nodes, n_kmers
  s.k = in.u64();
  if (s.k == 0 || s.k > 32) throw std::runtime_error("Unsupported k-mer size " + std::to_string(s.k));
  return s;
}

bool suffix_groups_consistent(const uint64_t *const acgt[4], const uint64_t *sgs, uint64_t n) {
  const uint64_t words = (n + 63) / 64;
  for (uint64_t w = 0; w < words; ++w) {
    uint64_t edges = acgt[0][w] | acgt[1][w] | acgt[2][w] | acgt[3][w];
    uint64_t starts = sgs[w];
    if (w == 0) starts |= 1;
    if (w + 1 == words && (n & 63)) edges &= (1ULL << (n & 63)) - 1;
    if (edges & ~starts) return false;
  }
  return true;
}

HostBitVector load_key_kmer_marks(const std::string &path) {
  In in(path);
  check_colors_magic(in);  // SbwtBuilder.cpp:147-165, skip per :167-182
  in.skip_bit_vector();    // dense_arrays
  in.skip_int_vector();    // dense_arrays_intervals
  in.skip_int_vector();    // sparse_arrays
  in.skip_int_vector();    // sparse_arrays_intervals
  in.skip_bit_vector();    // is_dense_marks
  in.skip_bit_vector();    // its rank_support_v5
  HostBitVector b;
  in.bit_vector(b);
  return b;
}

HostColors load_colors(const std::string &path) {
  In in(path);
  check_colors_magic(in);
  HostColors c;
  in.bit_vector(c.dense_arrays);
  in.int_vector(c.dense_arrays_intervals);
  in.int_vector(c.sparse_arrays);
  in.int_vector(c.sparse_arrays_intervals);
  in.bit_vector(c.is_dense_marks);
  in.skip_bit_vector();
  in.bit_vector(c.key_kmer_marks);
  in.skip_bit_vector();
  in.int_vector(c.color_set_idxs);
  c.num_color_sets = c.is_dense_marks.bits;
  in.u64();                      // max color set index
  c.num_colors = in.u64() + 1;   // ColorIndexBuilder.cpp:70-72
  in.u64();                      // cumulative sum of color set sizes
  return c;
}

}  // namespace sbwt_amd
