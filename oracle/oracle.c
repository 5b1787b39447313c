/*
 * oracle.c -- CPU restatement of SBWT-Search's search path (test infrastructure
 * only: see oracle.h). Plain C99 + OpenMP; every function names the reference
 * file:line it follows (paths relative to the SBWT-Search source tree).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HYPERBLOCK_BITS (1ULL << 32) /* src/Global/GlobalDefinitions.h.in:14 */
#define SUPERBLOCK_BITS 1024ULL      /* GlobalDefinitions.h.in:15 */
#define BASICBLOCK_BITS 256ULL       /* GlobalDefinitions.h.in:16 */
#define PAD_WORDS 8ULL               /* zero words kept after every bit vector */

static uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

/* ------------------------------------------------------------------------ */
/* Poppy: src/PoppyBuilder/PoppyBuilder.cpp:20-70.                           */
/* The reference emits one layer_1_2 entry per 1024-bit superblock up to     */
/* round_up(n, 1024); its rank reads entry n/1024 when n % 1024 == 0         */
/* (src/UtilityKernels/Rank.cuh:56). We run the same builder over one extra  */
/* all-zero superblock so that rank(n) is defined and equals total_1s.       */
/* ------------------------------------------------------------------------ */
int orc_poppy_build(const uint64_t *bits, uint64_t num_bits, orc_poppy *out) {
  const uint64_t end_bits = round_up(num_bits, SUPERBLOCK_BITS) + SUPERBLOCK_BITS;
  const uint64_t data_bits = round_up(num_bits, 64);
  out->n_layer_0 = end_bits / HYPERBLOCK_BITS + 1;
  out->n_layer_1_2 = end_bits / SUPERBLOCK_BITS;
  out->layer_0 = (uint64_t *)calloc(out->n_layer_0 + 1, 8);
  out->layer_1_2 = (uint64_t *)calloc(out->n_layer_1_2 + 1, 8);
  if (!out->layer_0 || !out->layer_1_2) return -1;
  uint64_t l0_count = 0, l1_count = 0, l2_count = 0;
  uint64_t l2_temps[3] = {0, 0, 0};
  uint64_t l2_idx = 0, n_l0 = 0, n_l12 = 0;
  for (uint64_t i = 0, b = 0; b < end_bits; b += 64, ++i) {
    if (b % SUPERBLOCK_BITS == 0) {
      if (b % HYPERBLOCK_BITS == 0) {
        out->layer_0[n_l0++] = l0_count;
        l1_count = 0;
      }
      l2_idx = 0;
      l2_count = 0;
      l2_temps[0] = l2_temps[1] = l2_temps[2] = 0;
    }
    if (b % BASICBLOCK_BITS == 0 && b % SUPERBLOCK_BITS != 0) {
      l2_temps[l2_idx++] = l2_count;
      l2_count = 0;
      if (l2_idx == 3) {
        out->layer_1_2[n_l12++] = ((l1_count - l2_temps[0] - l2_temps[1] - l2_temps[2]) << 32) |
                                  (l2_temps[0] << 20) | (l2_temps[1] << 10) | l2_temps[2];
      }
    }
    if (b < data_bits) {
      const uint64_t c = (uint64_t)__builtin_popcountll(bits[i]);
      l0_count += c;
      l1_count += c;
      l2_count += c;
    }
  }
  out->n_layer_0 = n_l0;
  out->n_layer_1_2 = n_l12;
  out->total_1s = l0_count;
  return 0;
}

void orc_poppy_free(orc_poppy *p) {
  free(p->layer_0);
  free(p->layer_1_2);
  p->layer_0 = p->layer_1_2 = NULL;
}

/* rank over [0, index): src/UtilityKernels/Rank.cuh:26-65 (bits LSB-first). */
uint64_t orc_rank(const uint64_t *bits, const orc_poppy *p, uint64_t index) {
  const uint64_t bb_first_word = (index / BASICBLOCK_BITS) * (BASICBLOCK_BITS / 64);
  const uint64_t in_bb_word = (index / 64) % (BASICBLOCK_BITS / 64);
  uint64_t basic = 0;
  for (uint64_t w = 0; w < in_bb_word; ++w) basic += (uint64_t)__builtin_popcountll(bits[bb_first_word + w]);
  const uint64_t r = index % 64;
  if (r) basic += (uint64_t)__builtin_popcountll(bits[bb_first_word + in_bb_word] << (64 - r));
  const uint64_t e12 = p->layer_1_2[index / SUPERBLOCK_BITS];
  const uint64_t joined = (e12 & ((1ULL << 30) - 1)) >> (10 * (3 - (index % SUPERBLOCK_BITS) / BASICBLOCK_BITS));
  const uint64_t l2 = (joined >> 20) + ((joined >> 10) & 1023) + (joined & 1023);
  return basic + l2 + (e12 >> 32) + p->layer_0[index / HYPERBLOCK_BITS];
}

uint64_t orc_rank_brute(const uint64_t *bits, uint64_t index) {
  uint64_t c = 0;
  for (uint64_t w = 0; w < index / 64; ++w) c += (uint64_t)__builtin_popcountll(bits[w]);
  if (index % 64) c += (uint64_t)__builtin_popcountll(bits[index / 64] & ((1ULL << (index % 64)) - 1));
  return c;
}

/* ------------------------------------------------------------------------ */
/* .sbwt plain-matrix v0.1 reader: src/SbwtBuilder/SbwtBuilder.cpp:58-145.   */
/* ------------------------------------------------------------------------ */
static int read_u64(FILE *f, uint64_t *v) { return fread(v, 8, 1, f) == 1 ? 0 : -1; }

static int read_sized_string(FILE *f, char *buf, size_t cap) {
  uint64_t len;
  if (read_u64(f, &len) || len >= cap) return -1;
  if (fread(buf, 1, len, f) != len) return -1;
  buf[len] = 0;
  return 0;
}

static int skip_bytes(FILE *f, uint64_t n) { return fseeko(f, (off_t)n, SEEK_CUR); }

static uint64_t *read_bit_words(FILE *f, uint64_t bits) {
  const uint64_t words = round_up(bits, 64) / 64;
  uint64_t *w = (uint64_t *)calloc(words + PAD_WORDS, 8);
  if (!w) return NULL;
  if (fread(w, 8, words, f) != words) {
    free(w);
    return NULL;
  }
  return w;
}

static void sbwt_finish(orc_sbwt *s) {
  s->c_map[0] = 1; /* SbwtBuilder.cpp:111 (cmap initialised to 1) */
  for (int c = 0; c < 4; ++c) {
    orc_poppy_build(s->acgt[c], s->num_bits, &s->poppy[c]);
    s->c_map[c + 1] = s->poppy[c].total_1s;
  }
  for (int c = 0; c < 4; ++c) s->c_map[c + 1] += s->c_map[c]; /* SbwtBuilder.cpp:130 */
}

int orc_sbwt_load(const char *path, orc_sbwt *s) {
  memset(s, 0, sizeof(*s));
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char buf[64];
  int rc = -1;
  if (read_sized_string(f, buf, sizeof buf)) goto done;
  if (strcmp(buf, "v0.1") != 0) { /* SbwtBuilder.cpp:61-68 */
    if (strcmp(buf, "plain-matrix") != 0) { rc = -2; goto done; }
    if (read_sized_string(f, buf, sizeof buf) || strcmp(buf, "v0.1") != 0) { rc = -3; goto done; }
  }
  uint64_t bits;
  for (int c = 0; c < 4; ++c) {
    if (read_u64(f, &bits)) goto done;
    if (c == 0) s->num_bits = bits;
    else if (bits != s->num_bits) goto done;
    if (!(s->acgt[c] = read_bit_words(f, bits))) goto done;
  }
  s->words = round_up(s->num_bits, 64) / 64;
  for (int i = 0; i < 4; ++i) { /* 4 rank structures: SbwtBuilder.cpp:97 */
    if (read_u64(f, &bits) || skip_bytes(f, round_up(bits, 64) / 8)) goto done;
  }
  if (read_u64(f, &bits)) goto done; /* suffix group starts */
  if (!(s->sgs = read_bit_words(f, bits))) goto done;
  uint64_t nbytes;
  for (int i = 0; i < 2; ++i) { /* C map, kmer_prefix_calc: SbwtBuilder.cpp:99-100 */
    if (read_u64(f, &nbytes) || skip_bytes(f, nbytes)) goto done;
  }
  if (skip_bytes(f, 24)) goto done; /* precalc_k, n_nodes, n_kmers */
  if (read_u64(f, &s->k)) goto done;
  sbwt_finish(s);
  rc = 0;
done:
  fclose(f);
  if (rc) orc_sbwt_free(s);
  return rc;
}

int orc_sbwt_from_bits(const uint64_t *const acgt[4], uint64_t num_bits, uint64_t k, orc_sbwt *s) {
  memset(s, 0, sizeof(*s));
  s->num_bits = num_bits;
  s->k = k;
  s->words = round_up(num_bits, 64) / 64;
  for (int c = 0; c < 4; ++c) {
    s->acgt[c] = (uint64_t *)calloc(s->words + PAD_WORDS, 8);
    if (!s->acgt[c]) return -1;
    memcpy(s->acgt[c], acgt[c], s->words * 8);
  }
  sbwt_finish(s);
  return 0;
}

void orc_sbwt_free(orc_sbwt *s) {
  for (int c = 0; c < 4; ++c) {
    free(s->acgt[c]);
    orc_poppy_free(&s->poppy[c]);
  }
  free(s->sgs);
  memset(s, 0, sizeof(*s));
}

/* ------------------------------------------------------------------------ */
/* Presearch: src/Presearcher/Presearcher.cuh:19-41, generalised to kp chars */
/* (the reference fixes kp = 12, GlobalDefinitions.h.in:17). Prefix x holds  */
/* its first char in the top two of its 2*kp bits.                           */
/* ------------------------------------------------------------------------ */
static inline void sbwt_step(const orc_sbwt *s, uint32_t c, uint64_t *l, uint64_t *r) {
  *l = s->c_map[c] + orc_rank(s->acgt[c], &s->poppy[c], *l);
  *r = s->c_map[c] + orc_rank(s->acgt[c], &s->poppy[c], *r + 1) - 1;
}

void orc_presearch(const orc_sbwt *s, uint32_t kp, uint64_t *left, uint64_t *right) {
  const uint64_t n = 1ULL << (2 * kp);
#pragma omp parallel for schedule(static)
  for (uint64_t x = 0; x < n; ++x) {
    uint32_t c = (uint32_t)(x >> (2 * kp - 2)) & 3;
    uint64_t l = s->c_map[c], r = s->c_map[c + 1] - 1;
    for (int sh = 2 * (int)kp - 4; sh >= 0; sh -= 2) sbwt_step(s, (uint32_t)(x >> sh) & 3, &l, &r);
    left[x] = l;
    right[x] = r;
  }
}

static inline int get_bit(const uint64_t *bv, uint64_t i) { return (int)((bv[i / 64] >> (i % 64)) & 1); }

/* Key-k-mer move: src/IndexSearcher/IndexSearcher.cuh:60-70. The reference   */
/* loops forever on a row with no outgoing edge; the oracle stops there and   */
/* returns ORC_WALK_STUCK.                                                    */
static uint64_t key_walk(const orc_sbwt *s, const uint64_t *key_marks, uint64_t l, uint32_t *steps) {
  uint32_t st = 0;
  while (!get_bit(key_marks, l)) {
    int moved = 0;
    for (uint32_t c = 0; c < 4; ++c) {
      if (get_bit(s->acgt[c], l)) {
        l = s->c_map[c] + orc_rank(s->acgt[c], &s->poppy[c], l);
        moved = 1;
        break;
      }
    }
    ++st;
    if (!moved || st > s->num_bits) {
      l = ORC_WALK_STUCK;
      break;
    }
  }
  if (steps) *steps = st;
  return l;
}

/* Search: src/IndexSearcher/IndexSearcher.cuh:22-72. The 64-bit window at    */
/* char p (:37-42), presearch lookup of its top 2*kp bits (:43-47), then one  */
/* C[c] + rank step per remaining char (:48-55).                              */
void orc_search(const orc_sbwt *s, uint32_t kp, const uint64_t *left, const uint64_t *right,
                const uint64_t *packed, const uint64_t *positions, uint64_t n_pos,
                const uint64_t *key_marks, uint64_t *out, uint8_t *steps_out, uint32_t *walk_out,
                int threads) {
  const uint64_t k = s->k;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads > 0 ? threads : 1)
  for (uint64_t idx = 0; idx < n_pos; ++idx) {
    const uint64_t pos = positions[idx];
    const uint64_t bit = pos * 2;
    const uint64_t first = packed[bit / 64] << (bit % 64);
    const uint64_t second = (bit % 64) ? (packed[bit / 64 + 1] >> (64 - bit % 64)) : 0;
    const uint64_t window = first | second;
    const uint64_t pre = window >> (64 - 2 * kp);
    uint64_t l = left[pre], r = right[pre];
    uint8_t steps = 0, emptied = (l > r);
    for (uint64_t i = pos + kp; i < pos + k; ++i) {
      const uint32_t c = (uint32_t)(packed[i / 32] >> (62 - (i % 32) * 2)) & 3;
      sbwt_step(s, c, &l, &r);
      if (!emptied) {
        ++steps;
        emptied = (l > r);
      }
    }
    uint32_t walk = 0;
    uint64_t res;
    if (l > r) res = ORC_NOT_FOUND;
    else res = key_marks ? key_walk(s, key_marks, l, &walk) : l;
    out[idx] = res;
    if (steps_out) steps_out[idx] = steps;
    if (walk_out) walk_out[idx] = walk;
  }
}

uint64_t orc_search_codes(const orc_sbwt *s, const uint8_t *codes, uint64_t len) {
  uint64_t l = 0, r = s->num_bits - 1;
  for (uint64_t i = 0; i < len && l <= r; ++i) sbwt_step(s, codes[i], &l, &r);
  return l <= r ? l : ORC_NOT_FOUND;
}

/* ------------------------------------------------------------------------ */
/* Packing: src/SeqToBitsConverter/CharToBits.h:21-40 and                    */
/* ContinuousSeqToBitsConverter.cpp:90-107: A/a 0, C/c 1, G/g 2, T/t 3, char */
/* j -> word j/32 bits [63-2(j%32), 62-2(j%32)]; other bytes pack as 0 and   */
/* set invalid[j] = 1.                                                       */
/* ------------------------------------------------------------------------ */
static int code_of(unsigned char ch) {
  switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

void orc_pack(const char *chars, uint64_t n, uint64_t *words, uint8_t *invalid) {
  const uint64_t nw = round_up(n, 32) / 32;
#pragma omp parallel for schedule(static)
  for (uint64_t w = 0; w < nw; ++w) {
    uint64_t v = 0;
    for (uint64_t j = w * 32; j < n && j < w * 32 + 32; ++j) {
      const int c = code_of((unsigned char)chars[j]);
      if (c < 0) {
        if (invalid) invalid[j] = 1;
        continue;
      }
      if (invalid) invalid[j] = 0;
      v |= (uint64_t)c << (62 - 2 * (j % 32));
    }
    words[w] = v;
  }
}

/* Positions: src/PositionsBuilder/PositionsBuilder.cpp:15-60: every start    */
/* s..e-k of each read [s, e) with e - s >= k.                                */
uint64_t orc_positions(const uint64_t *read_ends, uint64_t n_reads, uint64_t k, uint64_t *positions) {
  uint64_t n = 0, start = 0;
  for (uint64_t r = 0; r < n_reads; ++r) {
    const uint64_t end = read_ends[r];
    for (uint64_t p = start; p + k <= end; ++p) {
      if (positions) positions[n] = p;
      ++n;
    }
    start = end;
  }
  return n;
}

/* Classified per-k-mer results, the contract of                              */
/* src/IndexResultsPrinter/ContinuousIndexResultsPrinter.hpp:146-308: a k-mer  */
/* with any invalid char is ORC_INVALID, else a not-found search result is    */
/* ORC_NOT_FOUND, else the node index.                                        */
uint64_t orc_index_reads(const orc_sbwt *s, uint32_t kp, const uint64_t *left, const uint64_t *right,
                         const char *chars, const uint64_t *read_ends, uint64_t n_reads,
                         const uint64_t *key_marks, uint64_t *out, int threads) {
  const uint64_t n_chars = n_reads ? read_ends[n_reads - 1] : 0;
  const uint64_t k = s->k;
  uint64_t *packed = (uint64_t *)calloc(round_up(n_chars, 32) / 32 + 2, 8);
  uint8_t *invalid = (uint8_t *)calloc(n_chars + 1, 1);
  const uint64_t n_pos = orc_positions(read_ends, n_reads, k, NULL);
  uint64_t *positions = (uint64_t *)malloc((n_pos + 1) * 8);
  orc_pack(chars, n_chars, packed, invalid);
  orc_positions(read_ends, n_reads, k, positions);
  orc_search(s, kp, left, right, packed, positions, n_pos, key_marks, out, NULL, NULL, threads);
  /* prefix count of invalid chars, so a window test is one subtraction */
  uint64_t *inv_prefix = (uint64_t *)malloc((n_chars + 1) * 8);
  inv_prefix[0] = 0;
  for (uint64_t i = 0; i < n_chars; ++i) inv_prefix[i + 1] = inv_prefix[i] + invalid[i];
  for (uint64_t i = 0; i < n_pos; ++i) {
    const uint64_t p = positions[i];
    if (inv_prefix[p + k] - inv_prefix[p]) out[i] = ORC_INVALID;
  }
  free(inv_prefix);
  free(positions);
  free(invalid);
  free(packed);
  return n_pos;
}

/* ------------------------------------------------------------------------ */
/* FASTA/FASTQ reader. The reference delegates to kseqpp_REad v1.6.0          */
/* (third party, absent here; call site                                       */
/* src/SequenceFileParser/ContinuousSequenceFileParser.cpp:125,152). Restated */
/* from kseq's published record grammar: skip to the first '>' or '@'; the    */
/* header is the rest of that line; sequence = concatenation of following     */
/* lines up to a line starting with '>', '@' or '+'; after '+', quality lines */
/* are consumed until they are at least as long as the sequence. A trailing   */
/* '\r' is kept as a sequence byte (and so is invalid).                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  char *d;
  uint64_t n, cap;
} cbuf;
static void cb_push(cbuf *b, char c) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 4096;
    b->d = (char *)realloc(b->d, b->cap);
  }
  b->d[b->n++] = c;
}
typedef struct {
  uint64_t *d;
  uint64_t n, cap;
} ubuf;
static void ub_push(ubuf *b, uint64_t v) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 1024;
    b->d = (uint64_t *)realloc(b->d, b->cap * 8);
  }
  b->d[b->n++] = v;
}

int orc_read_seqfile(const char *path, orc_reads *out) {
  memset(out, 0, sizeof(*out));
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  fseeko(f, 0, SEEK_END);
  const long long sz = ftello(f);
  fseeko(f, 0, SEEK_SET);
  char *d = (char *)malloc((size_t)sz + 1);
  if (fread(d, 1, (size_t)sz, f) != (size_t)sz) {
    fclose(f);
    free(d);
    return -1;
  }
  fclose(f);
  cbuf seq = {0};
  ubuf ends = {0};
  long long i = 0;
  while (i < sz && d[i] != '>' && d[i] != '@') ++i;
  while (i < sz) {
    ++i;                                   /* header marker */
    while (i < sz && d[i] != '\n') ++i;    /* header line */
    if (i < sz) ++i;
    uint64_t seq_len = 0;
    while (i < sz && d[i] != '>' && d[i] != '+' && d[i] != '@') {
      if (d[i] == '\n') { ++i; continue; }
      while (i < sz && d[i] != '\n') { cb_push(&seq, d[i++]); ++seq_len; }
      if (i < sz) ++i;
    }
    ub_push(&ends, seq.n);
    if (i < sz && d[i] == '+') {
      while (i < sz && d[i] != '\n') ++i;  /* '+' line */
      if (i < sz) ++i;
      uint64_t q = 0;
      while (i < sz && q < seq_len) {
        while (i < sz && d[i] != '\n') { ++i; ++q; }
        if (i < sz) ++i;
      }
      while (i < sz && d[i] != '>' && d[i] != '@') ++i;
    }
  }
  free(d);
  out->chars = seq.d ? seq.d : (char *)calloc(1, 1);
  out->n_chars = seq.n;
  out->ends = ends.d ? ends.d : (uint64_t *)calloc(1, 8);
  out->n_reads = ends.n;
  return 0;
}

void orc_reads_free(orc_reads *r) {
  free(r->chars);
  free(r->ends);
  memset(r, 0, sizeof(*r));
}

/* u64 length + bytes: src/Tools/IOUtils.cpp:44-56. */
static void write_sized_string(FILE *f, const char *s) {
  const uint64_t n = strlen(s);
  fwrite(&n, 8, 1, f);
  fwrite(s, 1, n, f);
}

/* Index-stage output: Ascii (v1.0, "-1"/"-2", ' ' between, '\n' per read),  */
/* Binary (v1.0, u64 with u64max/-1 not found, -2 invalid, -3 newline) and    */
/* Bool (v2.0, '0' found, '1' not found, '2' invalid): src/IndexResultsPrinter */
/* /{Ascii,Binary,Bool}ContinuousIndexResultsPrinter.cpp, headers per          */
/* ContinuousIndexResultsPrinter.hpp:336-346.                                  */
int orc_index_file(const orc_sbwt *s, const char *in_path, const char *out_path, int format,
                   int write_headers, const uint64_t *key_marks, int threads) {
  orc_reads rd;
  if (orc_read_seqfile(in_path, &rd)) return -1;
  const uint32_t kp = s->k < 12 ? (uint32_t)s->k : 12;
  const uint64_t np = 1ULL << (2 * kp);
  uint64_t *left = (uint64_t *)malloc(np * 8), *right = (uint64_t *)malloc(np * 8);
  orc_presearch(s, kp, left, right);
  const uint64_t n_pos = orc_positions(rd.ends, rd.n_reads, s->k, NULL);
  uint64_t *res = (uint64_t *)malloc((n_pos + 1) * 8);
  orc_index_reads(s, kp, left, right, rd.chars, rd.ends, rd.n_reads, key_marks, res, threads);
  FILE *f = fopen(out_path, "wb");
  if (!f) return -1;
  static const char *fmt_names[3] = {"ascii", "binary", "bool"};
  static const char *fmt_vers[3] = {"v1.0", "v1.0", "v2.0"};
  if (write_headers) {
    write_sized_string(f, fmt_names[format]);
    write_sized_string(f, fmt_vers[format]);
  }
  uint64_t ri = 0, start = 0;
  for (uint64_t r = 0; r < rd.n_reads; ++r) {
    const uint64_t end = rd.ends[r];
    const uint64_t nk = end - start >= s->k ? end - start - s->k + 1 : 0;
    for (uint64_t j = 0; j < nk; ++j, ++ri) {
      const uint64_t v = res[ri];
      if (format == 0) {
        if (j) fputc(' ', f);
        if (v == ORC_INVALID) fputs("-2", f);
        else if (v == ORC_NOT_FOUND) fputs("-1", f);
        else fprintf(f, "%llu", (unsigned long long)v);
      } else if (format == 1) {
        fwrite(&v, 8, 1, f);
      } else {
        fputc(v == ORC_INVALID ? '2' : v == ORC_NOT_FOUND ? '1' : '0', f);
      }
    }
    if (format == 1) {
      const uint64_t nl = ORC_NEWLINE;
      fwrite(&nl, 8, 1, f);
    } else {
      fputc('\n', f);
    }
    start = end;
  }
  fclose(f);
  free(res);
  free(left);
  free(right);
  orc_reads_free(&rd);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Color index: .tcolors "sdsl-hybrid-v4" reader,                            */
/* src/ColorIndexBuilder/ColorIndexBuilder.cpp:23-86. sdsl-lite (c32874c)    */
/* on-disk forms: bit_vector = u64 bits + ceil(bits/64) words; int_vector<> = */
/* u64 bits + u8 width + ceil(bits/64) words; rank_support_v5 = one           */
/* int_vector<64> (u64 bits + words), skipped.                                */
/* ------------------------------------------------------------------------ */
static int load_bitv(FILE *f, orc_bitv *b, int with_poppy) {
  if (read_u64(f, &b->bits)) return -1;
  if (!(b->data = read_bit_words(f, b->bits))) return -1;
  if (with_poppy) orc_poppy_build(b->data, b->bits, &b->poppy);
  return 0;
}

static int load_intv(FILE *f, orc_intv *v) {
  uint64_t bits;
  unsigned char width;
  if (read_u64(f, &bits) || fread(&width, 1, 1, f) != 1 || width == 0 || width > 64) return -1;
  v->width = width;
  v->size = bits / width;
  if (!(v->data = read_bit_words(f, bits))) return -1;
  return 0;
}

static int skip_rank_v5(FILE *f) {
  uint64_t bits;
  if (read_u64(f, &bits)) return -1;
  return skip_bytes(f, round_up(bits, 64) / 8);
}

int orc_colors_load(const char *path, orc_colors *c) {
  memset(c, 0, sizeof(*c));
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char buf[64];
  int rc = -1;
  if (read_sized_string(f, buf, sizeof buf) || strcmp(buf, "sdsl-hybrid-v4") != 0) { rc = -2; goto done; }
  if (load_bitv(f, &c->dense_arrays, 0)) goto done;
  if (load_intv(f, &c->dense_arrays_intervals)) goto done;
  if (load_intv(f, &c->sparse_arrays)) goto done;
  if (load_intv(f, &c->sparse_arrays_intervals)) goto done;
  if (load_bitv(f, &c->is_dense_marks, 1) || skip_rank_v5(f)) goto done;
  if (load_bitv(f, &c->key_kmer_marks, 1) || skip_rank_v5(f)) goto done;
  if (load_intv(f, &c->color_set_idxs)) goto done;
  uint64_t max_cs, max_color, cum;
  if (read_u64(f, &max_cs) || read_u64(f, &max_color) || read_u64(f, &cum)) goto done;
  c->num_color_sets = c->is_dense_marks.bits;
  c->num_colors = max_color + 1; /* ColorIndexBuilder.cpp:70-72 */
  rc = 0;
done:
  fclose(f);
  if (rc) orc_colors_free(c);
  return rc;
}

static void free_bitv(orc_bitv *b) {
  free(b->data);
  orc_poppy_free(&b->poppy);
}

void orc_colors_free(orc_colors *c) {
  free_bitv(&c->dense_arrays);
  free(c->dense_arrays_intervals.data);
  free(c->sparse_arrays.data);
  free(c->sparse_arrays_intervals.data);
  free_bitv(&c->is_dense_marks);
  free_bitv(&c->key_kmer_marks);
  free(c->color_set_idxs.data);
  memset(c, 0, sizeof(*c));
}

/* sdsl int_vector element: src/UtilityKernels/VariableLengthIntegerIndex.cuh:23-36. */
uint64_t orc_intv_get(const orc_intv *v, uint64_t i) {
  const uint64_t bit = i * v->width, w = bit / 64, off = bit % 64;
  uint64_t x = v->data[w] >> off;
  if (off + v->width > 64) x |= v->data[w + 1] << (64 - off);
  return v->width == 64 ? x : x & ((1ULL << v->width) - 1);
}

/* Color set of a key node, the per-lane part of                             */
/* src/ColorSearcher/ColorSearcher.cuh:96-131 and its helpers :208-318:       */
/* key-mark rank -> color_set_idx -> dense bitmap [start, end) or sparse ids. */
uint64_t orc_color_set_of(const orc_colors *c, uint64_t sbwt_idx, uint32_t *ids) {
  const uint64_t j = orc_rank(c->key_kmer_marks.data, &c->key_kmer_marks.poppy, sbwt_idx);
  const uint64_t cs = orc_intv_get(&c->color_set_idxs, j);
  const uint64_t dr = orc_rank(c->is_dense_marks.data, &c->is_dense_marks.poppy, cs);
  uint64_t n = 0;
  if (get_bit(c->is_dense_marks.data, cs)) {
    const uint64_t start = orc_intv_get(&c->dense_arrays_intervals, dr);
    const uint64_t end = orc_intv_get(&c->dense_arrays_intervals, dr + 1);
    for (uint64_t b = start; b < end; ++b)
      if (get_bit(c->dense_arrays.data, b)) ids[n++] = (uint32_t)(b - start);
  } else {
    const uint64_t si = cs - dr;
    const uint64_t start = orc_intv_get(&c->sparse_arrays_intervals, si);
    const uint64_t end = orc_intv_get(&c->sparse_arrays_intervals, si + 1);
    for (uint64_t a = start; a < end; ++a) ids[n++] = (uint32_t)orc_intv_get(&c->sparse_arrays, a);
  }
  return n;
}

/* Warp layout of the color stage input: src/IndexFileParser/IndexFileParser.cpp */
/* :36-74 and AsciiIndexFileParser.cpp:33-70. Only found indexes are kept,     */
/* each read padded with u64max to a multiple of `warp`; warp_intervals holds   */
/* cumulative warp counts of reads with >= 1 found index; colored_seq_id maps   */
/* every read to its interval slot (a read with none shares the next slot).     */
uint64_t orc_warp_layout(const uint64_t *results, const uint64_t *read_kmer_ends, uint64_t n_reads,
                         uint64_t warp, uint64_t *warped, uint64_t *warp_intervals,
                         uint64_t *n_intervals, uint64_t *colored_seq_id, uint64_t *found,
                         uint64_t *not_found, uint64_t *invalid) {
  uint64_t nw = 0, ni = 1, start = 0;
  warp_intervals[0] = 0;
  for (uint64_t r = 0; r < n_reads; ++r) {
    uint64_t f = 0, nf = 0, inv = 0;
    for (uint64_t i = start; i < read_kmer_ends[r]; ++i) {
      const uint64_t v = results[i];
      if (v == ORC_INVALID) ++inv;
      else if (v == ORC_NOT_FOUND) ++nf;
      else {
        warped[nw++] = v;
        ++f;
      }
    }
    while (nw % warp) warped[nw++] = UINT64_MAX;
    colored_seq_id[r] = ni - 1;
    if (warp_intervals[ni - 1] != nw / warp) warp_intervals[ni++] = nw / warp;
    found[r] = f;
    not_found[r] = nf;
    invalid[r] = inv;
    start = read_kmer_ends[r];
  }
  *n_intervals = ni;
  return nw;
}

/* Per-read counts: the ballot/popcount of ColorSearcher.cuh:175-205 summed   */
/* over a read's warps by ColorPostProcessor.cuh:21-43, i.e. for every found  */
/* index of the read, +1 on each color of its color set.                      */
void orc_color_counts(const orc_colors *c, const uint64_t *results, const uint64_t *read_kmer_ends,
                      uint64_t n_reads, uint64_t *counts, int threads) {
  const uint64_t nc = c->num_colors;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    uint32_t *ids = (uint32_t *)malloc((nc + 1) * 4);
#pragma omp for schedule(dynamic, 64)
    for (uint64_t r = 0; r < n_reads; ++r) {
      uint64_t *row = counts + r * nc;
      memset(row, 0, nc * 8);
      const uint64_t start = r ? read_kmer_ends[r - 1] : 0;
      for (uint64_t i = start; i < read_kmer_ends[r]; ++i) {
        const uint64_t v = results[i];
        if (v == ORC_INVALID || v == ORC_NOT_FOUND) continue;
        const uint64_t m = orc_color_set_of(c, v, ids);
        for (uint64_t j = 0; j < m; ++j) row[ids[j]]++;
      }
    }
    free(ids);
  }
}

/* src/ColorResultsPrinter/ContinuousColorResultsPrinter.hpp:243-246. */
uint64_t orc_color_min_hits(uint64_t found, uint64_t not_found, uint64_t invalid, double threshold,
                            int include_not_found, int include_invalid) {
  const uint64_t seq_size = found + (uint64_t)(include_not_found != 0) * not_found +
                            (uint64_t)(include_invalid != 0) * invalid;
  return (uint64_t)ceil((double)seq_size * threshold);
}

/* ------------------------------------------------------------------------ */
/* Index-result file bodies (the colors stage's input), after the two sized  */
/* header strings.                                                          */
/* Ascii: the char loop of src/IndexFileParser/AsciiIndexFileParser.cpp     */
/* :33-70 with parse_number (:91-98) and skip_until_next_whitespace         */
/* (:85-89). Chars are signed; "a number" is any char >= '0'. '-' consumes  */
/* the next char (whatever it is: '1' counts a not-found, '2' an invalid),   */
/* then every char >= '0', then one more char, which ends the read if it is */
/* '\n'. A number's terminator is read again by the loop. A read ends at    */
/* each '\n' met by the loop (IndexFileParser.cpp:36-40).                   */
/* Binary: u64 words (BinaryIndexFileParser.cpp:33-70): u64max not found,   */
/* u64max-1 invalid, u64max-2 end of read, anything else found; a partial   */
/* last word is not read.                                                   */
/* results[] gets, in file order, the found indexes and ORC_NOT_FOUND /     */
/* ORC_INVALID for the counted "-1" / "-2"; kmer_ends[r] = results before   */
/* the end of read r. Results after the last read end belong to no read.   */
/* Capacity: results <= n + 1, reads <= n + 1. Returns the reads.           */
/* ------------------------------------------------------------------------ */
uint64_t orc_parse_index_body(const unsigned char *body, uint64_t n, int format,
                              uint64_t *results, uint64_t *n_results, uint64_t *kmer_ends) {
  uint64_t nr = 0, nl = 0;
  if (format == 1) {
    for (uint64_t w = 0; w + 8 <= n; w += 8) {
      uint64_t v;
      memcpy(&v, body + w, 8);
      if (v == ORC_NEWLINE) kmer_ends[nl++] = nr;
      else results[nr++] = v;
    }
    *n_results = nr;
    return nl;
  }
  uint64_t p = 0;
  while (p < n) {
    signed char c = (signed char)body[p++];
    if (c == '-') {
      if (p < n) {
        const signed char d = (signed char)body[p++];
        if (d == '1') results[nr++] = ORC_NOT_FOUND;
        else if (d == '2') results[nr++] = ORC_INVALID;
      }
      c = 0;
      while (p < n && (c = (signed char)body[p++]) >= '0') c = 0;
      if (c == '\n') kmer_ends[nl++] = nr;
      continue;
    }
    if (c == '\n') {
      kmer_ends[nl++] = nr;
      continue;
    }
    if (c >= '0') {
      uint64_t v = (uint64_t)(c - '0');
      while (p < n && (signed char)body[p] >= '0') v = v * 10 + (uint64_t)((signed char)body[p++] - '0');
      results[nr++] = v;
    }
  }
  *n_results = nr;
  return nl;
}

/* Thresholded color ids of n_reads reads (ids[offsets[r] .. offsets[r+1]),   */
/* ascending) as the color printers write them (without file headers):      */
/* mode 0 ascii, src/ColorResultsPrinter/AsciiContinuousColorResultsPrinter */
/* .cpp:62-79 with ContinuousColorResultsPrinter.hpp:237-263 (' ' between   */
/* ids, '\n' per read); 1 binary, BinaryContinuousColorResultsPrinter.cpp   */
/* :51-62 (u64 per id, u64max per read); 2 csv, CsvContinuousColorResults-  */
/* Printer.cpp:40-50,78-110 (a row of num_colors "0"/"1" joined by ',' and  */
/* ended by '\n'). out may be NULL (size only). Returns the byte count.     */
uint64_t orc_color_lists_format(const uint64_t *offsets, const uint32_t *ids, uint64_t n_reads,
                                uint64_t num_colors, int mode, char *out) {
  uint64_t b = 0;
  char tmp[24];
  for (uint64_t r = 0; r < n_reads; ++r) {
    if (mode == 2) {
      if (out) {
        for (uint64_t c = 0; c < num_colors; ++c) {
          out[b + 2 * c] = '0';
          out[b + 2 * c + 1] = c + 1 == num_colors ? '\n' : ',';
        }
        for (uint64_t i = offsets[r]; i < offsets[r + 1]; ++i) out[b + 2 * (uint64_t)ids[i]] = '1';
      }
      b += 2 * num_colors;
      continue;
    }
    for (uint64_t i = offsets[r]; i < offsets[r + 1]; ++i) {
      if (mode == 1) {
        const uint64_t v = ids[i];
        if (out) memcpy(out + b, &v, 8);
        b += 8;
        continue;
      }
      if (i > offsets[r]) {
        if (out) out[b] = ' ';
        ++b;
      }
      const int m = snprintf(tmp, sizeof tmp, "%u", ids[i]);
      if (out) memcpy(out + b, tmp, (size_t)m);
      b += (uint64_t)m;
    }
    if (mode == 1) {
      const uint64_t end = UINT64_MAX;
      if (out) memcpy(out + b, &end, 8);
      b += 8;
    } else {
      if (out) out[b] = '\n';
      ++b;
    }
  }
  return b;
}
