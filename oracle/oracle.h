/*
 * oracle.h -- CPU restatement of SBWT-Search's index-search and color-search
 * path, used ONLY as the checker.
 *
 * THIS IS TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only to check or to
 * time the CPU beside the GPU. The product library (libsbwt_amd.so) never
 * links, loads or calls anything in this directory.
 *
 * Every function restates the reference file:line named in its comment
 * (paths relative to the SBWT-Search source tree). Pinning of each function
 * (fixtures, closed forms, brute force) is listed in tests/test_oracle_*.py.
 */
#ifndef SBWT_ORACLE_H
#define SBWT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NOT_FOUND UINT64_MAX              /* ascii "-1" */
#define ORC_INVALID (UINT64_MAX - 1)          /* ascii "-2" */
#define ORC_NEWLINE (UINT64_MAX - 2)          /* binary-format read end */
#define ORC_WALK_STUCK (UINT64_MAX - 3)       /* key walk hit a row with no edge */

/* Reference-layout Poppy rank directory (src/Poppy/Poppy.h:19-27). */
typedef struct {
  uint64_t *layer_0;
  uint64_t n_layer_0;
  uint64_t *layer_1_2;
  uint64_t n_layer_1_2;
  uint64_t total_1s;
} orc_poppy;

/* An SBWT as read from a plain-matrix v0.1 file (src/SbwtBuilder/SbwtBuilder.cpp:58-132). */
typedef struct {
  uint64_t num_bits;   /* n = number of SBWT nodes */
  uint64_t k;
  uint64_t words;      /* ceil(n / 64) */
  uint64_t *acgt[4];   /* words + padding words of zero */
  orc_poppy poppy[4];
  uint64_t c_map[5];
  uint64_t *sgs;       /* suffix-group starts (read for tests; unused by search) */
} orc_sbwt;

/* ---- SBWT loading, Poppy, rank ---- */
int orc_sbwt_load(const char *path, orc_sbwt *out);
/* Build from raw bit vectors already in memory (copies them). */
int orc_sbwt_from_bits(const uint64_t *const acgt[4], uint64_t num_bits, uint64_t k,
                       orc_sbwt *out);
void orc_sbwt_free(orc_sbwt *s);
int orc_poppy_build(const uint64_t *bits, uint64_t num_bits, orc_poppy *out);
void orc_poppy_free(orc_poppy *p);
uint64_t orc_rank(const uint64_t *bits, const orc_poppy *p, uint64_t index);
/* Brute-force popcount rank over [0, index), the closed form rank must equal. */
uint64_t orc_rank_brute(const uint64_t *bits, uint64_t index);

/* ---- presearch + search ---- */
void orc_presearch(const orc_sbwt *s, uint32_t kp, uint64_t *left, uint64_t *right);
/* One search over explicit positions; out[i] = node index or ORC_NOT_FOUND.
 * key_marks may be NULL (no key-k-mer move). steps_out (nullable) receives
 * the number of rank steps the search needs before the interval is empty
 * (or k - kp if it never empties); walk_out (nullable) the key-walk length. */
void orc_search(const orc_sbwt *s, uint32_t kp, const uint64_t *left, const uint64_t *right,
                const uint64_t *packed, const uint64_t *positions, uint64_t n_pos,
                const uint64_t *key_marks, uint64_t *out, uint8_t *steps_out,
                uint32_t *walk_out, int threads);
/* Direct search of one k-mer given as ACGT codes, no presearch table. */
uint64_t orc_search_codes(const orc_sbwt *s, const uint8_t *codes, uint64_t len);

/* ---- 2-bit packing + invalid flags, positions ---- */
void orc_pack(const char *chars, uint64_t n, uint64_t *words, uint8_t *invalid);
uint64_t orc_positions(const uint64_t *read_ends, uint64_t n_reads, uint64_t k,
                       uint64_t *positions);

/* Per-k-mer classified results for a batch of reads (invalid > not found > index). */
uint64_t orc_index_reads(const orc_sbwt *s, uint32_t kp, const uint64_t *left,
                         const uint64_t *right, const char *chars, const uint64_t *read_ends,
                         uint64_t n_reads, const uint64_t *key_marks, uint64_t *out,
                         int threads);

/* ---- FASTA/FASTQ + index-stage output files ---- */
typedef struct {
  char *chars;         /* all sequences concatenated */
  uint64_t n_chars;
  uint64_t *ends;      /* cumulative end offset of each read */
  uint64_t n_reads;
} orc_reads;
int orc_read_seqfile(const char *path, orc_reads *out);
void orc_reads_free(orc_reads *r);
/* format: 0 ascii, 1 binary, 2 bool. Returns 0 on success. */
int orc_index_file(const orc_sbwt *s, const char *in_path, const char *out_path, int format,
                   int write_headers, const uint64_t *key_marks, int threads);

/* ---- color index (.tcolors, sdsl-hybrid-v4) ---- */
typedef struct {
  uint64_t *data;      /* words (+1 zero padding word) */
  uint64_t size;       /* elements */
  uint32_t width;      /* bits per element */
} orc_intv;
typedef struct {
  uint64_t *data;      /* words (+ padding) */
  uint64_t bits;
  orc_poppy poppy;     /* built only where the reference builds one */
} orc_bitv;
typedef struct {
  orc_bitv dense_arrays;
  orc_intv dense_arrays_intervals;
  orc_intv sparse_arrays;
  orc_intv sparse_arrays_intervals;
  orc_bitv is_dense_marks;
  orc_bitv key_kmer_marks;
  orc_intv color_set_idxs;
  uint64_t num_color_sets;
  uint64_t num_colors;
} orc_colors;
int orc_colors_load(const char *path, orc_colors *out);
void orc_colors_free(orc_colors *c);
uint64_t orc_intv_get(const orc_intv *v, uint64_t i);
/* Color set of one key node: writes sorted color ids, returns count. */
uint64_t orc_color_set_of(const orc_colors *c, uint64_t sbwt_idx, uint32_t *ids);
/* Warp-padded layout of found indexes (src/IndexFileParser/IndexFileParser.cpp:36-74). */
uint64_t orc_warp_layout(const uint64_t *results, const uint64_t *read_kmer_ends,
                         uint64_t n_reads, uint64_t warp, uint64_t *warped,
                         uint64_t *warp_intervals, uint64_t *n_intervals,
                         uint64_t *colored_seq_id, uint64_t *found, uint64_t *not_found,
                         uint64_t *invalid);
/* Per-read color counts counts[read * num_colors + c], u64 (ColorSearcher.cpp:38-58). */
void orc_color_counts(const orc_colors *c, const uint64_t *results,
                      const uint64_t *read_kmer_ends, uint64_t n_reads, uint64_t *counts,
                      int threads);
/* Body of an ascii (format 0) or binary (format 1) index-result file ->
 * per-k-mer results (found index / ORC_NOT_FOUND / ORC_INVALID) and cumulative
 * result counts at each read end (AsciiIndexFileParser.cpp:33-98,
 * BinaryIndexFileParser.cpp:33-70). Returns the number of reads. */
uint64_t orc_parse_index_body(const unsigned char *body, uint64_t n, int format,
                              uint64_t *results, uint64_t *n_results, uint64_t *kmer_ends);
/* Color id lists -> printer bytes, mode 0 ascii / 1 binary / 2 csv (no file
 * headers); out may be NULL. Returns the byte count. */
uint64_t orc_color_lists_format(const uint64_t *offsets, const uint32_t *ids, uint64_t n_reads,
                                uint64_t num_colors, int mode, char *out);
/* ceil(seq_size * threshold) with the reference's double arithmetic. */
uint64_t orc_color_min_hits(uint64_t found, uint64_t not_found, uint64_t invalid,
                            double threshold, int include_not_found, int include_invalid);

#ifdef __cplusplus
}
#endif
#endif
