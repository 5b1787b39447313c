/*
 * sbwt_amd.h -- C ABI of the MI355X (gfx950) SBWT-Search engine.
 *
 * The library replaces the reference's GPU seam: the device containers, the
 * presearcher and the two searchers. Every entry point names the reference
 * interface it replaces (paths relative to the SBWT-Search source tree).
 *
 * Conventions (as in the reference, src/IndexSearcher/IndexSearcher.cpp and
 * src/Tools/GpuUtils.h):
 *  - Every function returns 0 on success or a negative status; the message of
 *    the last failure on the calling thread is sbwt_last_error(). HIP failures
 *    read "GPUassert: <hipGetErrorString> at <file>:<line>" (GpuUtils.h:20-29).
 *  - Index handles own their device memory and are shared read-only by any
 *    number of searchers (IndexSearchMain.cpp:308-316). A searcher owns one
 *    HIP stream and its batch buffers and is not thread-safe by itself.
 *  - Calls taking host buffers are synchronous on return (IndexSearcher.cu:63-65).
 *    Calls ending in _device take device pointers and only enqueue work on the
 *    searcher's stream; sbwt_searcher_sync waits for it.
 *  - Bit vectors are arrays of u64 words, bit i of the vector = bit i % 64 of
 *    word i / 64 (the sdsl-lite bit_vector layout of the .sbwt/.tcolors files).
 *  - Result sentinels: SBWT_NOT_FOUND (u64 max, printed "-1"), SBWT_INVALID
 *    (u64 max - 1, printed "-2"), as in BinaryContinuousIndexResultsPrinter.
 */
#ifndef SBWT_AMD_H
#define SBWT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBWT_NOT_FOUND UINT64_MAX
#define SBWT_INVALID (UINT64_MAX - 1)
#define SBWT_WALK_STUCK (UINT64_MAX - 3)

typedef struct sbwt_index sbwt_index;
typedef struct sbwt_searcher sbwt_searcher;

/* Message of the last failure on this thread ("" if none). */
const char *sbwt_last_error(void);

/* Free and total bytes of the current device.
 * Replaces gpu_utils::get_free_gpu_memory (src/Tools/GpuUtils.cu:7-12). */
int sbwt_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);

/* Random-line probe, the ceiling the search kernels are measured against (no
 * counterpart in the reference): dependent random 16-byte loads, one per
 * 128-byte line, over a table of the smallest power-of-two number of lines
 * that holds working_set_bytes (allocated and freed by the call), from
 * blocks_per_cu 256-lane workgroups per CU each following `steps` lines.
 * *ms = one timed pass after an untimed one; *lines = lines read in it;
 * *table_bytes (nullable) = the working set used. Synchronous. */
int sbwt_random_line_probe(uint64_t working_set_bytes, uint32_t blocks_per_cu, uint32_t steps,
                           float *ms, uint64_t *lines, uint64_t *table_bytes);

/* ---- SBWT index -------------------------------------------------------- */

/* Device index from the four ACGT bit vectors of an SBWT with num_bits nodes
 * (ceil(num_bits/64) words each). Builds the rank layout, the c-map and the
 * presearch table of presearch_letters letters, at most min(k, 16): 4^kp
 * intervals of 8 bytes (16 if num_bits >= 2^32). Longer prefixes cut SBWT
 * steps from every search; results do not depend on it. 0 chooses about
 * log4(num_bits) + 2 letters within [min(k, 12), min(k, 16)], the table
 * taking at most an eighth of device memory (the reference fixes 12).
 * key_kmer_marks (nullable, ceil(num_bits/64) words) enables the
 * key-k-mer move. suffix_group_starts (nullable, the .sbwt file's vector of
 * the same name) lets the streaming search step from nodes that do not start
 * their suffix group; without it such steps fall back to a full search.
 * Replaces GpuSbwtContainer(acgt, poppy, c_map, bits_total, bit_vector_size,
 * kmer_size, key_kmer_marks) (src/SbwtContainer/GpuSbwtContainer.h:33-41,
 * GpuSbwtContainer.cpp:16-52) plus Presearcher::presearch()
 * (src/Presearcher/Presearcher.h:26-32, Presearcher.cpp:18-30); the c-map is
 * computed as in SbwtBuilder.cpp:111,128-130. */
int sbwt_index_create(const uint64_t *const acgt[4], uint64_t num_bits, uint32_t kmer_size,
                      const uint64_t *key_kmer_marks, const uint64_t *suffix_group_starts,
                      uint32_t presearch_letters, sbwt_index **out);

/* Device index from a .sbwt/.tdbg plain-matrix v0.1 file and, if colors_path
 * is not null, the key-k-mer marks of a .tcolors sdsl-hybrid-v4 file.
 * Replaces SbwtBuilder(dbg, colors).get_cpu_sbwt()->to_gpu() + presearch
 * (src/SbwtBuilder/SbwtBuilder.cpp:41-182, IndexSearchMain.cpp:87-103).
 * Errors: "Error input is not a plain-matrix SBWT", "Error: wrong SBWT
 * version", "The colors file has an incorrect format. Expected
 * 'sdsl-hybrid-v4'" (SbwtBuilder.cpp:64,67,152-154). */
int sbwt_index_load(const char *sbwt_path, const char *colors_path, uint32_t presearch_letters,
                    sbwt_index **out);

/* 1 if suffix_group_starts fits the four ACGT vectors of an SBWT of num_bits
 * nodes, else 0: every node not marked as a group start has an empty row (in
 * an SBWT only the first node of a suffix group carries the group's edges,
 * which the reference's interval step C[c] + rank_c relies on,
 * src/IndexSearcher/IndexSearcher.cuh:48-55). sbwt_index_create and
 * sbwt_index_load run this check; when it fails the rows are not an SBWT's
 * and every search of that index runs per k-mer (the reference's interval
 * search on the file's rows), whatever algorithm a searcher asks for. Host
 * only, no device work. No counterpart in the reference, which never reads
 * the vector. */
int sbwt_suffix_groups_consistent(const uint64_t *const acgt[4], const uint64_t *suffix_group_starts,
                                  uint64_t num_bits);

/* Builds the index's k-mer table: every node whose label is a whole k-mer,
 * by its 2k-bit code, so that a search that cannot step from the previous
 * k-mer (a read start, a k-mer after a miss) reads about one 32-byte group of
 * entries instead of the presearch entry and k - presearch_letters SBWT
 * steps. Results do not change. One-time cost (about 0.1 s for 10^8 nodes)
 * and 32 to 64 bytes per node of device memory (4x that at k = 32), so it
 * pays for an index that serves many batches; the command line does not
 * build it. *built (nullable) receives 0 when the index cannot have one
 * (nodes >= 2^32, k > 32, presearch_letters >= k, or the table would take
 * more than half of device memory or leave less than an eighth free). No
 * counterpart in the reference (its searches always start from the
 * presearch table, src/IndexSearcher/IndexSearcher.cuh:36-55). */
int sbwt_index_build_kmer_table(sbwt_index *index, int *built);

void sbwt_index_destroy(sbwt_index *index);

typedef struct {
  uint64_t num_nodes;         /* SbwtContainer::get_num_bits */
  uint32_t kmer_size;         /* SbwtContainer::get_kmer_size */
  uint32_t presearch_letters;
  uint64_t c_map[5];
  uint64_t device_bytes;
  int has_key_kmer_marks;
  int wide;                   /* node indexes need 64 bits */
  int has_suffix_groups;      /* the streaming search steps through group starts */
  int suffix_groups_rejected; /* starts were given but failed sbwt_suffix_groups_consistent */
} sbwt_index_info;
int sbwt_index_get_info(const sbwt_index *index, sbwt_index_info *info);

/* ---- index search ------------------------------------------------------ */

/* A searcher for batches of up to max_chars_per_batch chars and
 * max_reads_per_batch reads. hip_stream (nullable) is the hipStream_t to run
 * on; null creates a stream owned by the searcher.
 * Replaces IndexSearcher(stream_id, container, max_chars_per_batch,
 * move_to_key_kmer) (src/IndexSearcher/IndexSearcher.h:34-39, IndexSearcher.cpp:14-24). */
int sbwt_searcher_create(sbwt_index *index, uint64_t max_chars_per_batch,
                         uint64_t max_reads_per_batch, int move_to_key_kmer, void *hip_stream,
                         sbwt_searcher **out);
void sbwt_searcher_destroy(sbwt_searcher *searcher);
int sbwt_searcher_sync(sbwt_searcher *searcher);

/* Reference interface: bit_seqs (2-bit MSB-first, 32 chars per word, as built
 * by ContinuousSeqToBitsConverter) and one start position per k-mer ->
 * results[i] = SBWT index of the k-mer at positions[i] (after the key-k-mer
 * move if enabled) or SBWT_NOT_FOUND. Host buffers; synchronous.
 * Replaces IndexSearcher::search(bit_seqs, kmer_positions, results, batch_id)
 * (src/IndexSearcher/IndexSearcher.h:41-46, IndexSearcher.cpp:26-51). */
int sbwt_searcher_search(sbwt_searcher *searcher, const uint64_t *bit_seqs, uint64_t n_words,
                         const uint64_t *kmer_positions, uint64_t n_positions, uint64_t *results);

/* Same with device pointers; enqueued on the searcher's stream. */
int sbwt_searcher_search_device(sbwt_searcher *searcher, const uint64_t *d_bit_seqs,
                                const uint64_t *d_kmer_positions, uint64_t n_positions,
                                uint64_t *d_results);

/* Fused batch: ASCII reads (all reads concatenated, n_chars bytes) and the
 * cumulative end offset of every read -> one classified result per k-mer, in
 * read order: SBWT_INVALID if any of its k chars is not ACGTacgt, else the
 * search result. Replaces, in one call, ContinuousSeqToBitsConverter
 * (src/SeqToBitsConverter/ContinuousSeqToBitsConverter.cpp:65-107),
 * PositionsBuilder (src/PositionsBuilder/PositionsBuilder.cpp:15-60),
 * IndexSearcher::search and the invalid/not-found classification of
 * ContinuousIndexResultsPrinter (src/IndexResultsPrinter/
 * ContinuousIndexResultsPrinter.hpp:146-308). Host buffers; synchronous;
 * *n_kmers receives the number of results written. */
int sbwt_searcher_search_reads(sbwt_searcher *searcher, const char *chars, uint64_t n_chars,
                               const uint64_t *read_ends, uint64_t n_reads, uint64_t *results,
                               uint64_t *n_kmers);

/* Same with device pointers; enqueued on the searcher's stream. d_results
 * must hold n_chars entries. n_kmers (nullable, host) waits for the count. */
int sbwt_searcher_search_reads_device(sbwt_searcher *searcher, const char *d_chars,
                                      uint64_t n_chars, const uint64_t *d_read_ends,
                                      uint64_t n_reads, uint64_t *d_results, uint64_t *n_kmers);

/* Allocation-free fused search: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The per-k-mer fused search
 * of sbwt_searcher_search_reads_device (same kernels, same results) of one
 * batch of at most SBWT_NOALLOC_MAX_CHARS chars and SBWT_NOALLOC_MAX_READS
 * reads, without a searcher: the packed sequence, invalid masks, k-mer ends
 * and scan sums live in a static device array of the library (about 3.7 MB
 * per GPU, held until the process ends) instead of a searcher's buffers, so
 * the call makes no hipMalloc, hipFree or other allocation on the device.
 * d_results receives one result per k-mer, in read order (as many as the
 * reads have k-mers: sum of max(0, length - k + 1)). Enqueued on hip_stream
 * (null: the null stream); the index's device must be current. Every call on
 * a GPU shares that GPU's array and it keeps the last call's contents: calls
 * that overlap on one GPU (other streams or threads) race, so serialize them.
 * Larger batches are rejected with an error. */
#define SBWT_NOALLOC_MAX_CHARS (1ULL << 22)
#define SBWT_NOALLOC_MAX_READS (1ULL << 18)
int sbwt_index_search_reads_noalloc(const sbwt_index *index, int move_to_key_kmer,
                                    const char *d_chars, uint64_t n_chars,
                                    const uint64_t *d_read_ends, uint64_t n_reads,
                                    uint64_t *d_results, void *hip_stream);

/* Device time of the last call's kernels (ms): packing, k-mer offsets,
 * search. The reference logs the same with GpuEvent pairs
 * (src/IndexSearcher/IndexSearcher.cu:22,62,66-70). */
int sbwt_searcher_last_timings(sbwt_searcher *searcher, float *pack_ms, float *offsets_ms,
                               float *search_ms);
/* Of the last call's search time, the key-k-mer runs pass (streaming search
 * with move_to_key_kmer; 0 otherwise). No counterpart in the reference, whose
 * key move sits inside its search kernel (IndexSearcher.cuh:60-70). */
int sbwt_searcher_last_key_runs_time(sbwt_searcher *searcher, float *ms);

/* Census mode for sbwt_searcher_search_reads(_device): when enabled, the next
 * calls also count the exact algorithmic work of the batch: stats[0] full
 * searches (one presearch line each), stats[1] their SBWT steps (one
 * 128-byte node line each), stats[2] key-walk steps (one node line each),
 * stats[3] invalid k-mers, stats[4] k-mers stepped from the previous k-mer's
 * node (one node line each), stats[5] extra lines of those steps (the suffix
 * group began in an earlier line), stats[6] of those steps that fell back to
 * a full search, stats[7] of those steps the ones from a node that does not
 * start its suffix group, stats[8] absence checks (one presearch line each:
 * a kp-letter substring of the k-mer whose interval is empty is in no k-mer
 * of the index, so every k-mer containing it is not found), stats[9] k-mers
 * settled by such a check, stats[10] further k-mers settled by an earlier
 * check without a lookup, stats[11] further 128-byte lines of k-mer table
 * probes (a probe run's first line counts as its full search's presearch
 * line; a group in a line the run already read adds none);
 * stats[12..15] are zero. Results are unchanged. No counterpart in the
 * reference (its kernels keep no counters). */
int sbwt_searcher_set_stats(sbwt_searcher *searcher, int enable);
int sbwt_searcher_last_stats(sbwt_searcher *searcher, uint64_t stats[16]);

/* Search algorithm of the reads path: SBWT_ALGO_STREAMING steps each k-mer of
 * a read from the previous k-mer's node and searches from the presearch table
 * only after a miss (throughput: about one line per found k-mer), in work
 * items of the launch shape's chunk_kmers; SBWT_ALGO_PER_KMER searches every
 * k-mer independently, as the reference's d_search does
 * (IndexSearcher.cuh:22-72) (latency: a small batch takes one search's
 * time); SBWT_ALGO_AUTO (default) searches per k-mer below 2^22 chars and
 * streams above, with work items shrunk (down to 8 k-mers) until every
 * persistent lane has one; SBWT_ALGO_STREAMING_ADAPTIVE streams with those
 * shrunk work items from 2^16 chars (per k-mer below), for batches that
 * share the GPU with others. All return identical results. No counterpart in
 * the reference. */
#define SBWT_ALGO_STREAMING 0
#define SBWT_ALGO_PER_KMER 1
#define SBWT_ALGO_AUTO 2
#define SBWT_ALGO_STREAMING_ADAPTIVE 3
int sbwt_searcher_set_algorithm(sbwt_searcher *searcher, int algorithm);

/* Launch shape of the streaming search: k-mers per work item (chunk_kmers in
 * [64, 65535], default 256) and persistent 256-lane workgroups per CU
 * (blocks_per_cu in [1, 8], default 4). Results do not depend on it. No
 * counterpart in the reference (its launch is one thread per k-mer). */
int sbwt_searcher_set_launch(sbwt_searcher *searcher, uint32_t chunk_kmers, uint32_t blocks_per_cu);
int sbwt_searcher_get_launch(sbwt_searcher *searcher, uint32_t *chunk_kmers, uint32_t *blocks_per_cu);

/* Autotuner: runs the streaming search of one caller batch (device buffers,
 * as sbwt_searcher_search_reads_device) for each shape of a small grid
 * (chunk_kmers 128..1024 x blocks_per_cu 2..8), keeps the fastest search time
 * and returns it (outputs nullable). d_results holds the batch's results,
 * identical for every shape. Synchronous. No counterpart in the reference. */
int sbwt_searcher_autotune(sbwt_searcher *searcher, const char *d_chars, uint64_t n_chars,
                           const uint64_t *d_read_ends, uint64_t n_reads, uint64_t *d_results,
                           uint32_t *chunk_kmers, uint32_t *blocks_per_cu, float *search_ms);

/* Device 2-bit packing alone: chars -> ceil(n/32)+2 words (MSB-first) and
 * ceil(n/32)+2 u32 invalid masks (bit j%32 of word j/32 set if char j is not
 * ACGTacgt). Replaces ContinuousSeqToBitsConverter::parallel_generate
 * (ContinuousSeqToBitsConverter.cpp:65-107). hip_stream nullable. */
int sbwt_pack_2bit_device(const char *d_chars, uint64_t n_chars, uint64_t *d_words,
                          uint32_t *d_invalid, void *hip_stream);

/* Page-locked host memory (hipHostMalloc) for the host side of batch copies:
 * copies from and to it run at full link speed. The reference keeps its host
 * batches in PinnedVector (src/Tools/PinnedVector.h). */
int sbwt_host_alloc(uint64_t bytes, void **out);
void sbwt_host_free(void *ptr);

/* ---- Device formatting of index results ------------------------------ */

/* Formats index results into the bytes of the reference's index output files
 * (headers excluded; write them first): ascii "v v v\n" with -1 not found and
 * -2 invalid, binary u64 per result and u64 max - 2 per read end, bool '0'
 * found / '1' not found / '2' invalid and '\n' per read end. Replaces the
 * per-result loops of ContinuousIndexResultsPrinter::process_batch
 * (src/ResultsPrinter/ContinuousIndexResultsPrinter.hpp:146-308) and the
 * Ascii/Binary/Bool printers' converters (AsciiContinuousIndexResultsPrinter.cpp:50-85,
 * BinaryContinuousIndexResultsPrinter.cpp:52-78, BoolContinuousIndexResultsPrinter.cpp:43-69).
 * hip_stream nullable (the formatter then owns one). */
#define SBWT_FORMAT_ASCII 0
#define SBWT_FORMAT_BINARY 1
#define SBWT_FORMAT_BOOL 2
typedef struct sbwt_formatter sbwt_formatter;
int sbwt_formatter_create(uint64_t max_kmers, uint64_t max_pieces, void *hip_stream,
                          sbwt_formatter **out);
void sbwt_formatter_destroy(sbwt_formatter *formatter);

/* One batch of read pieces: piece e holds the results
 * d_results[d_piece_kmer_ends[e-1] .. d_piece_kmer_ends[e]) (cumulative, non-
 * decreasing; sbwt_searcher_kmer_ends_device gives them for a search batch)
 * and ends its read when d_piece_closes[e] != 0, else the next piece continues
 * the same line; line_open: the first piece continues a line on which the
 * previous batch already wrote results. *n_bytes receives the batch's byte count; the bytes go to
 * d_out only when d_out is not null and capacity >= *n_bytes (a call with
 * d_out = NULL sizes the buffer). Synchronous. */
int sbwt_formatter_format(sbwt_formatter *formatter, int mode, const uint64_t *d_results,
                          const uint64_t *d_piece_kmer_ends, const uint8_t *d_piece_closes,
                          uint64_t n_pieces, int line_open, char *d_out, uint64_t capacity,
                          uint64_t *n_bytes);

/* Host batch in, formatted index output bytes out: the search of
 * sbwt_searcher_search_reads (results stay on the device), then the formatter
 * on its pieces = the batch's reads (read_closes[e] != 0 when read e is
 * complete; 0 when the next read, in this batch or the next, continues its
 * line), one copy back. *n_bytes receives the byte count; the bytes are
 * copied to out when capacity >= *n_bytes, else they wait in the formatter for
 * sbwt_formatter_copy_last. Synchronous. Replaces one batch of the
 * reference's ContinuousIndexSearcher + ContinuousIndexResultsPrinter
 * (src/IndexSearcher/ContinuousIndexSearcher.cpp:67-74,
 * ContinuousIndexResultsPrinter.hpp:146-308). */
int sbwt_searcher_search_reads_format(sbwt_searcher *searcher, sbwt_formatter *formatter, int mode,
                                      const char *chars, uint64_t n_chars,
                                      const uint64_t *read_ends, const uint8_t *read_closes,
                                      uint64_t n_reads, int line_open, char *out,
                                      uint64_t capacity, uint64_t *n_bytes);
int sbwt_formatter_copy_last(sbwt_formatter *formatter, char *out, uint64_t capacity);

/* Device pointer to the cumulative k-mer count of every read of the last
 * sbwt_searcher_search_reads(_device) call (n_reads entries), valid until the
 * next call; it chains the index stage into the color stage on device. No
 * counterpart in the reference (its stages meet through files). */
int sbwt_searcher_kmer_ends_device(sbwt_searcher *searcher, const uint64_t **d_kmer_ends);

/* ---- color index and color search -------------------------------------- */

typedef struct sbwt_color_index sbwt_color_index;
typedef struct sbwt_color_searcher sbwt_color_searcher;

/* Device color index from a .tcolors sdsl-hybrid-v4 file.
 * Replaces ColorIndexBuilder(filename).get_cpu_color_index_container() and
 * CpuColorIndexContainer::to_gpu() -> GpuColorIndexContainer
 * (src/ColorIndexBuilder/ColorIndexBuilder.cpp:13-86,
 * src/ColorIndexContainer/GpuColorIndexContainer.cpp:13-86). */
int sbwt_color_index_load(const char *tcolors_path, sbwt_color_index **out);
void sbwt_color_index_destroy(sbwt_color_index *colors);

typedef struct {
  uint64_t num_colors;       /* GpuColorIndexContainer::num_colors (max color + 1) */
  uint64_t num_color_sets;   /* GpuColorIndexContainer::num_color_sets */
  uint64_t num_nodes;        /* length of the key-k-mer marks */
  uint64_t device_bytes;
} sbwt_color_index_info;
int sbwt_color_index_get_info(const sbwt_color_index *colors, sbwt_color_index_info *info);

/* Replaces ColorSearcher(stream_id, container, max_indexes_per_batch,
 * max_seqs_per_batch) (src/ColorSearcher/ColorSearcher.h:36-41,
 * ColorSearcher.cpp:16-36). hip_stream as in sbwt_searcher_create; give both
 * searchers the same stream to chain the stages on the device. */
int sbwt_color_searcher_create(sbwt_color_index *colors, uint64_t max_indexes_per_batch,
                               uint64_t max_reads_per_batch, void *hip_stream,
                               sbwt_color_searcher **out);
void sbwt_color_searcher_destroy(sbwt_color_searcher *searcher);

/* Reference interface: found SBWT (key) indexes, each read padded with
 * u64 max to a multiple of 64, and the cumulative number of 64-index warps of
 * every read with at least one found index (n_intervals = reads + 1) ->
 * results[read * num_colors + color] = number of the read's indexes whose
 * color set holds the color. Host buffers; synchronous.
 * Replaces ColorSearcher::search(sbwt_index_idxs, warps_intervals, results,
 * batch_id) (src/ColorSearcher/ColorSearcher.h:43-48, ColorSearcher.cpp:38-58). */
int sbwt_color_searcher_search(sbwt_color_searcher *searcher, const uint64_t *warped_idxs,
                               uint64_t n_idxs, const uint64_t *warp_intervals,
                               uint64_t n_intervals, uint64_t *results);

/* Threshold of the color stage: a read reports color c when its count of c is
 * at least ceil(seq_size * threshold) with seq_size = found
 * + include_not_found * not_found + include_invalid * invalid
 * (src/ColorResultsPrinter/ContinuousColorResultsPrinter.hpp:235-263).
 * first_read_quirk reproduces the reference printer's initial carry of u64 max
 * (ContinuousColorResultsPrinter.hpp:62-64,178-180): the first read of a
 * stream is thresholded with each of its three counts minus one (mod 2^64). */
typedef struct {
  double threshold;
  int include_not_found;
  int include_invalid;
  int first_read_quirk;
} sbwt_color_params;

/* Fused pseudoalignment: the per-k-mer results of sbwt_searcher_search_reads
 * (with the key-k-mer move) and the cumulative k-mer count of every read ->
 * for read r the sorted color ids ids[offsets[r] .. offsets[r+1]) passing the
 * threshold, and read_stats[3r..3r+2] = {found, not found, invalid}
 * (nullable). Replaces ColorSearcher::search + d_post_process + the
 * threshold of ContinuousColorResultsPrinter::do_print_seq in one pass, and
 * writes a sorted u32 list per read instead of reads x num_colors u64 counts.
 * Host buffers; synchronous; *n_ids = ids written. Fails if more than
 * ids_capacity ids are needed. */
int sbwt_color_searcher_pseudoalign(sbwt_color_searcher *searcher, const uint64_t *kmer_results,
                                    const uint64_t *read_kmer_ends, uint64_t n_reads,
                                    const sbwt_color_params *params, uint64_t *offsets,
                                    uint32_t *ids, uint64_t ids_capacity, uint32_t *read_stats,
                                    uint64_t *n_ids);

/* Same on device buffers (d_offsets: n_reads + 1 entries); synchronous on the
 * id count only. */
int sbwt_color_searcher_pseudoalign_device(sbwt_color_searcher *searcher,
                                           const uint64_t *d_kmer_results,
                                           const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                           const sbwt_color_params *params, uint64_t *d_offsets,
                                           uint32_t *d_ids, uint64_t ids_capacity,
                                           uint32_t *d_read_stats, uint64_t *n_ids);

/* Raw per-read counts on device: d_counts[read * num_colors + color] (u32). */
int sbwt_color_searcher_counts_device(sbwt_color_searcher *searcher,
                                      const uint64_t *d_kmer_results,
                                      const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                      uint32_t *d_counts);

/* Work distribution of the fused color search (sbwt_color_searcher_pseudoalign*):
 * item_kmers = 0 (default) gives one wave per read; otherwise the batch runs
 * over a device queue of (read, k-mer range) work items of item_kmers k-mers
 * (at least 64; lengthened when the partial-count slots of long reads would
 * take more than a 32nd of device memory), so reads of very uneven length
 * spread over many waves. Results are identical. No counterpart in the
 * reference (its color kernel takes one thread per index,
 * src/ColorSearcher/ColorSearcher.cuh:96-205). */
int sbwt_color_searcher_set_work_items(sbwt_color_searcher *searcher, uint32_t item_kmers);

/* Autotuner of the work-item shape above for one caller batch (the arguments
 * of sbwt_color_searcher_pseudoalign_device, whose outputs it overwrites): one
 * wave per read, then items of 256, 512, 1024, 2048 and 4096 k-mers, each
 * timed on one pass after an untimed one; the fastest is kept. *item_kmers
 * and *ms (nullable) receive it and its color-stage time. No counterpart in
 * the reference. */
int sbwt_color_searcher_autotune(sbwt_color_searcher *searcher, const uint64_t *d_kmer_results,
                                 const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                 const sbwt_color_params *params, uint64_t *d_offsets,
                                 uint32_t *d_ids, uint64_t ids_capacity, uint32_t *item_kmers,
                                 float *ms);

/* Fused index + color path (no counterpart in the reference, whose stages
 * meet through index files): builds the index's color table from a color
 * index over the same SBWT (the index needs its key-k-mer marks): per node,
 * its key k-mer's color set as a 64-bit mask when the index has at most 61
 * colors (*kind = 1; 8 bytes per node), else its key's rank among the key
 * k-mers (*kind = 2; 4 bytes per node). A searcher with color output
 * (sbwt_searcher_set_color_output) then writes each found k-mer's table entry
 * instead of its key node (one table line per run of k-mers sharing a key,
 * replacing the key move), and a color searcher whose input is that index
 * (sbwt_color_searcher_set_input) counts masks by ballots with no lookups, or
 * ranks with one key-set lookup per run instead of the key-mark rank and the
 * key set. Results are identical to the two-stage path. kind nullable. */
int sbwt_index_build_color_table(sbwt_index *index, const sbwt_color_index *colors, int *kind);
/* on: SBWT_COLOR_OUTPUT_ENTRIES (1) writes the table entries above;
 * SBWT_COLOR_OUTPUT_CHAINS (2, indexes of fewer than 2^32 nodes, reads path)
 * writes each found k-mer's node with what the search learnt about the k-mer
 * before it (its key's color entry when that one is marked), leaving the key
 * pass to a color searcher with chain input (sbwt_color_searcher_set_chain_input):
 * one pass over the results fewer; SBWT_COLOR_OUTPUT_NONE (0) key nodes. */
#define SBWT_COLOR_OUTPUT_NONE 0
#define SBWT_COLOR_OUTPUT_ENTRIES 1
#define SBWT_COLOR_OUTPUT_CHAINS 2
int sbwt_searcher_set_color_output(sbwt_searcher *searcher, int on);
/* index: the index whose color-table entries the results are (null: key nodes,
 * the default). */
int sbwt_color_searcher_set_input(sbwt_color_searcher *searcher, const sbwt_index *index);
/* The results are the key chains a searcher with SBWT_COLOR_OUTPUT_CHAINS
 * writes over index: the color stage resolves each chain to its key's entry
 * (as the searcher's key pass would) while it counts. Lists are identical to
 * the other inputs'. No counterpart in the reference. */
int sbwt_color_searcher_set_chain_input(sbwt_color_searcher *searcher, const sbwt_index *index);

/* Device time (ms) of the last call's color kernel. */
int sbwt_color_searcher_last_time(sbwt_color_searcher *searcher, float *ms);

/* Census of the fused color stage (measurement; results unchanged; no
 * counterpart in the reference). With enable set, each
 * sbwt_color_searcher_pseudoalign* call counts its exact algorithmic work:
 * census[0] key-set entries read (one 128-byte line each: a run of k-mers
 * sharing a key reads one), census[1] key-mark entries read (key-node input:
 * one line each), census[2] 128-byte lines spanned by the payloads of the
 * decoded sets, census[3] k-mer results read (8 bytes each). last_census
 * synchronizes the searcher's stream. */
int sbwt_color_searcher_set_census(sbwt_color_searcher *searcher, int enable);
int sbwt_color_searcher_last_census(sbwt_color_searcher *searcher, uint64_t census[4]);

/* ---- colors stage from index-result files --------------------------------- */

/* Bodies of the index stage's output files (after the two sized header
 * strings) parsed on the device into the color search's input: per-k-mer
 * results (found index, u64 max for a counted "-1", u64 max - 1 for "-2") and
 * cumulative result counts at each read end. Replaces the reference's
 * AsciiIndexFileParser / BinaryIndexFileParser::generate_batch
 * (src/IndexFileParser/AsciiIndexFileParser.cpp:33-98,
 * BinaryIndexFileParser.cpp:33-70) and the warp layout of
 * IndexFileParser.cpp:36-74, which the fused color search does not need.
 * Ascii follows the reference's char loop: any char >= '0' (signed) is a
 * digit; '-' takes the next char ('1' counts not found, '2' invalid), skips
 * the digits after it and consumes one more char, which ends the read if it
 * is '\n'; any other '\n' ends a read. Binary: u64 words, u64 max - 2 ends a
 * read, a partial last word is not read. Results after the last read end
 * belong to no read. max_bytes_per_batch < 2^32. */
#define SBWT_INDEX_FILE_ASCII 0
#define SBWT_INDEX_FILE_BINARY 1
typedef struct sbwt_index_file_parser sbwt_index_file_parser;
int sbwt_index_file_parser_create(uint64_t max_bytes_per_batch, uint64_t max_reads_per_batch,
                                  void *hip_stream, sbwt_index_file_parser **out);
void sbwt_index_file_parser_destroy(sbwt_index_file_parser *parser);
/* d_body: device bytes, 16-byte aligned (read in whole 16-byte chunks).
 * *d_results / *d_read_kmer_ends point into the parser (valid until its next
 * call); synchronous on the counts. */
int sbwt_index_file_parser_parse_device(sbwt_index_file_parser *parser, int format,
                                        const char *d_body, uint64_t n_bytes,
                                        const uint64_t **d_results,
                                        const uint64_t **d_read_kmer_ends, uint64_t *n_results,
                                        uint64_t *n_reads);
/* Host bytes in, host arrays out (capacities in entries); synchronous. */
int sbwt_index_file_parser_parse(sbwt_index_file_parser *parser, int format, const char *body,
                                 uint64_t n_bytes, uint64_t *results, uint64_t results_capacity,
                                 uint64_t *read_kmer_ends, uint64_t reads_capacity,
                                 uint64_t *n_results, uint64_t *n_reads);
/* Longest prefix of body[0, n_bytes) that a batch may end with: it ends right
 * after a read end ('\n' not preceded by '-' in ascii; a u64 max - 2 word in
 * binary) and holds at most max_reads read ends. 0 if there is none. Host
 * only, no device work. The reference's parser instead stops anywhere and
 * carries the open read into the next batch
 * (ContinuousColorResultsPrinter.hpp:175-185). */
uint64_t sbwt_index_file_batch_cut(int format, const char *body, uint64_t n_bytes,
                                   uint64_t max_reads);
/* Longest prefix of body[0, n_bytes) (starting at a read start) that ends
 * right after the read end of a read with at least one found index; 0 if
 * none. Cutting a batch there keeps each read without a found index in one
 * batch with the next read that has one: with --include-not-found or
 * --include-invalid the reference prints such a read with that read's counts
 * (colored_seq_id, src/IndexFileParser/IndexFileParser.cpp:65-74,
 * ContinuousColorResultsPrinter.hpp:201-209). Host only, no device work. */
uint64_t sbwt_index_file_colored_cut(int format, const char *body, uint64_t n_bytes);

/* Sequence-file chunks parsed on the device, the input side of the index
 * command: raw FASTA bytes that start at a record's header line and end at a
 * record's end -> the concatenated sequence chars and the cumulative end of
 * every read, the batch form that sbwt_searcher_search_reads takes. The
 * grammar is the reference's record reader (kseq++ behind
 * src/SequenceFileParser/ContinuousSequenceFileParser.cpp:72-112,148-156)
 * on input with no line starting with '+': a line starting with '>' or '@'
 * starts a record (its header is skipped), every other line is sequence whose
 * bytes other than '\n' are chars. A chunk holding a line that starts with '+'
 * (FASTQ quality) is rejected; such input stays on a host reader.
 * max_bytes_per_chunk < 2^32. */
typedef struct sbwt_fasta_parser sbwt_fasta_parser;
int sbwt_fasta_parser_create(uint64_t max_bytes_per_chunk, uint64_t max_reads_per_chunk,
                             void *hip_stream, sbwt_fasta_parser **out);
void sbwt_fasta_parser_destroy(sbwt_fasta_parser *parser);
/* Host bytes in, host chars and read ends out (capacities in entries);
 * synchronous. Errors when raw does not start with '>' or '@', holds a line
 * starting with '+', or more than max_reads_per_chunk records. */
int sbwt_fasta_parser_parse(sbwt_fasta_parser *parser, const char *raw, uint64_t n_bytes,
                            char *chars, uint64_t chars_capacity, uint64_t *read_ends,
                            uint64_t reads_capacity, uint64_t *n_chars, uint64_t *n_reads);
/* Longest prefix of raw[0, n_bytes) that a chunk may end with: it ends right
 * before a line starting with '>' or '@' (or is all of raw when last != 0,
 * the end of the file) and holds at most max_reads records. 0 if there is
 * none (one record is longer than raw). Host only, no device work. */
uint64_t sbwt_fasta_chunk_cut(const char *raw, uint64_t n_bytes, int last, uint64_t max_reads);
/* 1 if raw[0, n_bytes) holds a line starting with '+' (its first byte
 * counts when at_line_start != 0). Host only. */
int sbwt_fasta_has_quality_line(const char *raw, uint64_t n_bytes, int at_line_start);
/* One FASTA chunk -> device parse -> fused search -> formatted index output
 * bytes (every read complete: the chunk ends at a record's end). *n_bytes and
 * the copy-out follow sbwt_searcher_search_reads_format; *n_reads receives
 * the records of the chunk. Synchronous. Replaces one batch of the
 * reference's ContinuousSequenceFileParser + ContinuousIndexSearcher +
 * ContinuousIndexResultsPrinter (ContinuousSequenceFileParser.cpp:72-112,
 * src/IndexSearcher/ContinuousIndexSearcher.cpp:67-74,
 * ContinuousIndexResultsPrinter.hpp:146-308). */
int sbwt_searcher_search_fasta_format(sbwt_searcher *searcher, sbwt_fasta_parser *parser,
                                      sbwt_formatter *formatter, int mode, const char *raw,
                                      uint64_t n_raw, char *out, uint64_t capacity,
                                      uint64_t *n_bytes, uint64_t *n_reads);

/* FASTQ chunks in four-line records (header, one sequence line, '+' line, one
 * quality line) parsed on the device. The reference's record reader (same
 * citation as above) reads such a record as its sequence line exactly when
 * the header line starts with '@' or '>', the sequence line is not empty and
 * does not start with '+', '@' or '>', the third line starts with '+' and the
 * quality line is at least as long as the sequence line; the parser checks
 * this per record. raw starts at a record's header line; last != 0 when raw
 * ends the file (a last line without '
' then counts). The parser takes the
 * longest prefix of at most max_reads_per_chunk whole records that pass:
 * *n_used receives its length in bytes (0: no such record, e.g. one record
 * longer than raw or the first one failing), *n_chars / *n_reads its sequence
 * chars and reads. What follows the prefix is the caller's: the next chunk,
 * or the host reader when the prefix stopped at a failing record. chars /
 * read_ends may be NULL (the parse then stays on the device for
 * sbwt_searcher_search_fastq_format); otherwise capacities are in entries.
 * Synchronous. max_bytes_per_chunk < 2^32, max_reads_per_chunk < 2^30. */
typedef struct sbwt_fastq_parser sbwt_fastq_parser;
int sbwt_fastq_parser_create(uint64_t max_bytes_per_chunk, uint64_t max_reads_per_chunk,
                             void *hip_stream, sbwt_fastq_parser **out);
void sbwt_fastq_parser_destroy(sbwt_fastq_parser *parser);
int sbwt_fastq_parser_parse(sbwt_fastq_parser *parser, const char *raw, uint64_t n_bytes, int last,
                            char *chars, uint64_t chars_capacity, uint64_t *read_ends,
                            uint64_t reads_capacity, uint64_t *n_used, uint64_t *n_chars,
                            uint64_t *n_reads);
/* The parser's last parse -> fused search -> formatted index output bytes
 * (every read complete), as sbwt_searcher_search_fasta_format. Synchronous.
 * Replaces one batch of ContinuousIndexSearcher + ContinuousIndexResultsPrinter
 * (src/IndexSearcher/ContinuousIndexSearcher.cpp:67-74,
 * ContinuousIndexResultsPrinter.hpp:146-308) over a parsed FASTQ chunk. */
int sbwt_searcher_search_fastq_format(sbwt_searcher *searcher, sbwt_fastq_parser *parser,
                                      sbwt_formatter *formatter, int mode, char *out,
                                      uint64_t capacity, uint64_t *n_bytes);

/* Color id lists (the CSR of sbwt_color_searcher_pseudoalign_device) as the
 * reference's color printers write them, file headers excluded: ascii ids
 * joined by ' ' and a '\n' per read; binary a u64 per id and u64 max per read;
 * csv a row of num_colors "0"/"1" joined by ',' and ended by '\n'
 * (src/ColorResultsPrinter/AsciiContinuousColorResultsPrinter.cpp:62-79,
 * BinaryContinuousColorResultsPrinter.cpp:51-62,
 * CsvContinuousColorResultsPrinter.cpp:40-50,78-110). */
#define SBWT_COLOR_FORMAT_ASCII 0
#define SBWT_COLOR_FORMAT_BINARY 1
#define SBWT_COLOR_FORMAT_CSV 2
typedef struct sbwt_color_formatter sbwt_color_formatter;
int sbwt_color_formatter_create(uint64_t max_ids, uint64_t max_reads, void *hip_stream,
                                sbwt_color_formatter **out);
void sbwt_color_formatter_destroy(sbwt_color_formatter *formatter);
/* *n_bytes receives the byte count; the bytes go to d_out only when d_out is
 * not null and capacity >= *n_bytes. Synchronous on the count. */
int sbwt_color_formatter_format(sbwt_color_formatter *formatter, int mode,
                                const uint64_t *d_offsets, const uint32_t *d_ids,
                                uint64_t n_reads, uint64_t num_colors, char *d_out,
                                uint64_t capacity, uint64_t *n_bytes);

/* One batch of the colors command on one stream: index-file body bytes (host)
 * -> parse -> color search -> printer bytes (host). *n_bytes receives the
 * byte count (copied to out when capacity suffices, else kept for
 * sbwt_color_file_searcher_copy_last), *n_reads the reads of the batch. The
 * batch must end at a read end (sbwt_index_file_batch_cut). params->
 * first_read_quirk belongs on a stream's first batch only. Replaces one batch
 * of the reference's ContinuousIndexFileParser + ContinuousColorSearcher +
 * ContinuousColorResultsPrinter (ContinuousIndexFileParser.cpp:48-120,
 * ContinuousColorSearcher.cpp:73-80, ContinuousColorResultsPrinter.hpp
 * :167-263). */
typedef struct sbwt_color_file_searcher sbwt_color_file_searcher;
int sbwt_color_file_searcher_create(sbwt_color_index *colors, uint64_t max_bytes_per_batch,
                                    uint64_t max_reads_per_batch, void *hip_stream,
                                    sbwt_color_file_searcher **out);
void sbwt_color_file_searcher_destroy(sbwt_color_file_searcher *searcher);
int sbwt_color_file_searcher_batch(sbwt_color_file_searcher *searcher, int in_format,
                                   const char *body, uint64_t n_bytes,
                                   const sbwt_color_params *params, int out_mode, char *out,
                                   uint64_t capacity, uint64_t *n_bytes_out, uint64_t *n_reads);
int sbwt_color_file_searcher_copy_last(sbwt_color_file_searcher *searcher, char *out,
                                       uint64_t capacity);

/* ---- serving loop -------------------------------------------------------- */

/* A request loop over resident indexes: submitted batches of reads queue up
 * and are admitted to worker slots (each its own HIP stream, searchers and
 * pinned staging) as slots free up, so requests that arrive while others run
 * overlap on the device. The number of slots is min(slots, what
 * memory_fraction of the free device memory holds). Requests larger than a
 * slot run as several batches in that slot. Batches of up to 2^20 chars
 * replay a HIP graph captured once per slot and power-of-two (chars, reads)
 * bucket (staging copies, search, color stage, copies back in one launch;
 * no_graphs = 1 launches them one by one). No counterpart in the reference,
 * which processes its input files in fixed streams (IndexSearchMain.cpp:414-448). */
typedef struct sbwt_server sbwt_server;
typedef struct {
  uint32_t slots;
  uint64_t max_chars_per_slot;
  uint64_t max_reads_per_slot;
  double memory_fraction;
  int move_to_key_kmer;
  sbwt_color_params color;   /* used when a color index is attached */
  int no_graphs;
  int no_kmer_results;  /* with a color index: return only color lists (n_kmers = 0),
                           not the 8 bytes per k-mer of index results */
} sbwt_server_config;

typedef struct {
  uint64_t n_kmers;
  const uint64_t *kmer_results;   /* as sbwt_searcher_search_reads */
  uint64_t n_reads;
  const uint64_t *color_offsets;  /* n_reads + 1 entries, null without colors */
  const uint32_t *color_ids;
  double submit_s, start_s, end_s; /* seconds since the server started */
} sbwt_server_result;

int sbwt_server_create(sbwt_index *index, sbwt_color_index *colors /* nullable */,
                       const sbwt_server_config *config, sbwt_server **out);
/* Copies the request and returns at once with its id. */
int sbwt_server_submit(sbwt_server *server, const char *chars, uint64_t n_chars,
                       const uint64_t *read_ends, uint64_t n_reads, uint64_t *request_id);
/* Waits for the request; the result's buffers stay valid until release. A
 * failed request (e.g. malformed read ends) returns its error; other requests
 * are unaffected. */
int sbwt_server_wait(sbwt_server *server, uint64_t request_id, sbwt_server_result *out);
int sbwt_server_release(sbwt_server *server, uint64_t request_id);
/* Load generator over the serving loop, for measurement (no counterpart in
 * the reference): request i is reads [first_read[i], first_read[i+1]) of one
 * concatenated batch (chars and their cumulative read_ends; first_read has
 * n_requests + 1 entries, from 0), submitted without copying from the calling
 * thread at arrival_s[i] seconds (non-decreasing) after the call starts, while
 * a second thread waits for the requests in order and releases them. Per
 * request: latency_s[i] = its end minus its scheduled arrival (time spent
 * waiting behind a late submitter counts, so the schedule is never
 * coordinated away), lag_s[i] = its submit minus its scheduled arrival, and
 * digest[i] (nullable) a 64-bit hash of its k-mer results and color lists.
 * *wall_s receives the call's duration. Returns the first failed request's
 * error after all requests have ended. */
int sbwt_server_replay(sbwt_server *server, const char *chars, const uint64_t *read_ends,
                       const uint64_t *first_read, uint64_t n_requests, const double *arrival_s,
                       double *latency_s, double *lag_s, uint64_t *digest, double *wall_s);
int sbwt_server_get_slots(sbwt_server *server, uint32_t *slots, uint64_t *peak_in_flight);
void sbwt_server_destroy(sbwt_server *server);

#ifdef __cplusplus
}
#endif
#endif
