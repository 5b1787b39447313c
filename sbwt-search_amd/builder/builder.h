/*
 * builder.h -- synthetic workloads and index construction for SBWT-Search's
 * search path: seeded genomes and reads, an SBWT (plain-matrix v0.1) builder
 * and writer, and a color index (.tcolors, sdsl-hybrid-v4) builder and writer.
 *
 * These produce the inputs the search path consumes; nothing here sits inside a
 * timed region. bld_gpu_build constructs the same SBWT and color index on the
 * GPU for collections of billions of bases (gpu_build.hip).
 *
 * Random numbers: counter-based splitmix64 (Steele, Lea, Flood 2014),
 *   mix64(z) = splitmix64 finalizer of z + 0x9E3779B97F4A7C15,
 *   draw(seed, stream, i) = mix64(mix64(seed * 0xD1B54A32D192ED03 + stream) + i),
 * so every genome, read and base is reproducible in parallel.
 */
#ifndef SBWT_BUILDER_H
#define SBWT_BUILDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t bld_draw(uint64_t seed, uint64_t stream, uint64_t i);

/* i.i.d. uniform ACGT codes (0..3). */
void bld_genome_random(uint64_t seed, uint64_t len, uint8_t *out);
/* Copy of `root` with i.i.d. substitutions at `sub_rate` and indels at
 * `indel_rate` per base (length uniform in [1, max_indel], insertion or deletion
 * with equal odds). Writes at most out_cap codes, returns the length. */
uint64_t bld_genome_mutate(const uint8_t *root, uint64_t len, uint64_t seed, double sub_rate,
                           double indel_rate, uint32_t max_indel, uint8_t *out, uint64_t out_cap);

/* Reads of fixed or per-read lengths, as ASCII. For read r of length lens[r]:
 * with probability `random_frac` uniform random ACGT; else a window of genome
 * g (uniform among n_genomes) at a uniform start on a uniform strand, with
 * i.i.d. substitutions at `err_rate`. Then each base becomes 'N' with
 * probability n_rate. ends[r] = cumulative end offset. Returns total chars.
 * If mode_alternate != 0, even reads are sampled and odd reads random (the
 * exact half/half split of the 1 Mbp workload). */
uint64_t bld_reads(uint64_t seed, const uint8_t *genomes, const uint64_t *genome_ends,
                   uint64_t n_genomes, uint64_t n_reads, const uint32_t *lens, double random_frac,
                   double err_rate, double n_rate, int mode_alternate, char *out, uint64_t *ends,
                   uint32_t *src_genome);

/* A seeded uniform permutation of 0..n-1 (Fisher-Yates). */
void bld_permutation(uint64_t seed, uint64_t n, uint64_t *out);

/* Log-normal read lengths exp(N(log(median), sigma)) clipped to [lo, hi]. */
void bld_lognormal_lengths(uint64_t seed, uint64_t n, double median, double sigma, uint32_t lo,
                           uint32_t hi, uint32_t *out);
void bld_uniform_lengths(uint64_t seed, uint64_t n, uint32_t lo, uint32_t hi, uint32_t *out);

/* ---- SBWT ---- */
typedef struct bld_sbwt bld_sbwt;
/* Build from sequences (codes 0..3, concatenated; seq_ends cumulative). With
 * add_rc != 0 the reverse complement of every k-mer is added. k in [2, 32]. */
bld_sbwt *bld_sbwt_build(const uint8_t *codes, const uint64_t *seq_ends, uint64_t n_seqs,
                         uint32_t k, int add_rc, int threads);
void bld_sbwt_free(bld_sbwt *s);
uint64_t bld_sbwt_num_nodes(const bld_sbwt *s);
uint64_t bld_sbwt_num_kmers(const bld_sbwt *s);
/* Copies ceil(n/64) words of bit vector c (0..3 = A,C,G,T; 4 = suffix-group starts). */
void bld_sbwt_bits(const bld_sbwt *s, int c, uint64_t *out);
int bld_sbwt_write(const bld_sbwt *s, const char *path);
/* Node index of each k-mer (colex key, first char in the low bits) or UINT64_MAX.
 * Returns -1 (and writes nothing) when the structure holds no k-mer list (a GPU
 * build without keep_kmers). */
int bld_sbwt_lookup(const bld_sbwt *s, const uint64_t *keys, uint64_t n, uint64_t *out);

/* ---- color index ---- */
/* Colors: genome g of the sequences given to bld_sbwt_build (with add_rc a
 * genome holds a k-mer if it holds it or its reverse complement) maps to color
 * color_of[g]. Key k-mers: every node that is not alone in its suffix group,
 * has out-degree != 1, precedes a node of another color set, or is every d-th
 * node on the remaining non-branching paths (so the forward walk to a key
 * k-mer takes at most d - 1 steps). A color set is stored dense when its
 * num_colors-bit map is smaller than its id list. */
typedef struct bld_colors bld_colors;
bld_colors *bld_colors_build(const bld_sbwt *s, const uint8_t *codes, const uint64_t *seq_ends,
                             uint64_t n_seqs, const uint32_t *color_of, uint32_t d, int threads);
void bld_colors_free(bld_colors *c);
uint64_t bld_colors_num_colors(const bld_colors *c);
uint64_t bld_colors_num_sets(const bld_colors *c);
uint64_t bld_colors_num_dense(const bld_colors *c);
/* Copies the key-k-mer marks (ceil(n/64) words). */
void bld_colors_key_marks(const bld_colors *c, uint64_t *out);
int bld_colors_write(const bld_colors *c, const char *path);

/* ---- GPU construction (current HIP device) ----
 * The SBWT of bld_sbwt_build and, when color_of is given, the color index of
 * bld_colors_build with the same d, bit for bit: same nodes, bit vectors,
 * C map, color sets, set ids, key marks and set-id vector. k in [2, 31];
 * color ids below 2^31. Device memory: about 16 bytes per k-mer occurrence for
 * the sort (24 with colors) plus one byte per occurrence for flags; 1000
 * genomes of 2-5 Mbp with reverse complements (7 G occurrences) take about
 * 190 GB. keep_kmers != 0 keeps the sorted k-mer keys on the host (8 bytes per
 * k-mer), which bld_sbwt_lookup and bld_colors_build need.
 * Two distinct color lists with one 64-bit list hash are not supported (an
 * error, at about m^2 / 2^65 odds for m sets). Returns NULL on failure with
 * the reason in err; the stats hold what was counted up to the failure. */
typedef struct bld_gpu_stats {
  uint64_t occurrences, pairs, kmers, missing_prefixes, dummies, nodes, sets, key_nodes,
      cycle_nodes;
  uint64_t peak_device_bytes;
  double ms_keys, ms_sort, ms_sbwt, ms_colors, ms_total;
} bld_gpu_stats;
bld_sbwt *bld_gpu_build(const uint8_t *codes, const uint64_t *seq_ends, uint64_t n_seqs,
                        uint32_t k, int add_rc, const uint32_t *color_of, uint32_t d,
                        int keep_kmers, bld_colors **colors, bld_gpu_stats *stats, char *err,
                        uint64_t err_len);

#ifdef __cplusplus
}
#endif
#endif
