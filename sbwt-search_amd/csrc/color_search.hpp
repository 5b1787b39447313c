// Color search (pseudoalignment) on the MI355X: device color index and the
// ColorSearcher that replaces src/ColorSearcher/ColorSearcher.{h,cpp,cu}.
#pragma once

#include "device_index.hpp"
#include "loaders.hpp"

namespace sbwt_amd {

// One 16-byte entry per color set: {start lo, start hi, length, dense flag}.
// Dense: bits [start, start+length) of dense_arrays, color c present iff bit
// start+c. Sparse: sorted ids sparse_arrays[start .. start+length).
// This fuses the reference's is_dense bit, is_dense rank and the two interval
// reads (src/ColorSearcher/ColorSearcher.cuh:111-131, 208-290) into one line.
struct ColorView {
  const u64 *key_marks;       // mark entries {count, 64 bits} of the key-k-mer marks
  const u64 *color_set_idxs;  // sdsl int_vector words (+ padding)
  const uint4 *set_table;
  const uint4 *key_sets;      // set-table entry of every key k-mer, by key rank
  const u64 *dense_arrays;
  const u64 *sparse_arrays;
  u64 num_nodes;
  u64 num_color_sets;
  u32 csi_width;
  u32 sparse_width;
  u32 num_colors;
  // key-chain input (a searcher's color output kColorOutputChains): the
  // index's color table (per node) and, with masks, its key masks by rank
  const void *chain_table;
  const u64 *chain_key_masks;
};

struct DeviceColors {
  DeviceMarks key_marks;
  u64 *color_set_idxs = nullptr;
  uint4 *set_table = nullptr;
  uint4 *key_sets = nullptr;  // [key rank, 0..num_keys] = set_table[color_set_idxs[key rank]] (len 0 if none)
  u64 num_keys = 0;
  u64 *dense_arrays = nullptr;
  u64 *sparse_arrays = nullptr;
  u64 num_nodes = 0, num_color_sets = 0, num_colors = 0;
  u32 csi_width = 0, sparse_width = 0;
  u64 device_bytes = 0;
  void build(const HostColors &c);
  void release();
  ColorView view() const;
};

struct ColorParams {
  double threshold = 1.0;     // tau, --threshold (ColorSearchArgumentParser.cpp:126-148)
  bool include_not_found = false;
  bool include_invalid = false;
  // The reference's printer starts its carried per-read counts at u64 max
  // (ContinuousColorResultsPrinter.hpp:62-64,178-180), so the first read of a
  // stream is thresholded with found-1, not_found-1, invalid-1 (mod 2^64).
  bool first_read_quirk = false;
  // Quirk 2 needs no flag: with include_not_found or include_invalid, a read
  // with no found index is thresholded against the counts of the next read of
  // the batch that has one (IndexFileParser.cpp:65-74).
};

// The index's color table from a color index over the same SBWT: per node,
// its key k-mer's color set as a mask (at most 61 colors) or its key's rank
// among the key k-mers (more colors). Replaces any table built before.
void build_color_table(DeviceIndex &index, const DeviceColors &colors);

class ColorSearcher {
 public:
  ColorSearcher(const DeviceColors *colors, u64 max_indexes, u64 max_reads, hipStream_t stream);
  ~ColorSearcher();

  // Reference interface: warp-padded found indexes + cumulative warp counts
  // per colored read -> u64 counts[(n_intervals - 1) * num_colors]
  // (ColorSearcher::search, ColorSearcher.cpp:38-58). Host buffers, synchronous.
  void search_warped_host(const u64 *warped, u64 n_idxs, const u64 *warp_intervals,
                          u64 n_intervals, u64 *counts);

  // Fused form on device buffers: per-k-mer index results (index /
  // kNotFound / kInvalid, after the key-k-mer move) and cumulative k-mer
  // counts per read -> per-read thresholded color ids, sorted, as a CSR
  // (d_offsets[n_reads + 1], d_ids), plus per-read {found, not_found, invalid}.
  // Returns the total number of ids (synchronous on the count).
  u64 pseudoalign_device(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                         const ColorParams &p, u64 *d_offsets, u32 *d_ids, u64 ids_capacity,
                         u32 *d_stats);
  // The same without the host sync: nothing checks capacity on the host (ids
  // past it are not written) and the total is d_offsets[n_reads] once the
  // stream gets there. Capturable into a HIP graph (with timing off).
  void pseudoalign_device_async(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                                const ColorParams &p, u64 *d_offsets, u32 *d_ids, u64 ids_capacity,
                                u32 *d_stats);
  // Counts only (no threshold): u32 counts[n_reads * num_colors] on device.
  void counts_device(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads, u32 *d_counts);

  hipStream_t stream() const { return stream_; }
  float last_search_ms();
  void set_timing(bool on) { timing_ = on; }
  // Work distribution of the fused color search: 0 (default) one wave per
  // read; otherwise a device queue of (read, k-mer range) items of item_kmers
  // k-mers (lengthened when the partial-count slots of long reads would take
  // more than a 32nd of device memory), the batched form for reads of uneven
  // length. Results are identical.
  void set_work_items(u32 item_kmers);
  u32 work_items() const { return item_kmers_; }
  // Work-item shape for a caller batch (the arguments of pseudoalign_device):
  // one wave per read, then items of 256 to 4096 k-mers, each timed on one
  // pseudoalignment after an untimed one; keeps and returns the fastest
  // (*best_ms, nullable: its color-stage time). Outputs are overwritten.
  u32 autotune_items(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads, const ColorParams &p,
                     u64 *d_offsets, u32 *d_ids, u64 ids_capacity, float *best_ms);
  // Input of the fused color stage: the key nodes of the index search (null,
  // default), or the color-table entries that a searcher with color output
  // writes over index (whose table must come from this color index), or, with
  // chains, the key chains such a searcher writes with kColorOutputChains
  // (narrow indexes), resolved here instead of by the searcher's key pass.
  void set_input(const DeviceIndex *index, bool chains = false);
  ColorView input_view() const;  // the color index's view, with the chain input's tables
  // Census of the fused color stage (measurement; leaves results unchanged):
  // when on, each pseudoalign call counts {key-set entries read, key-mark
  // entries read, 128-byte payload lines spanned by the decoded sets, k-mers},
  // read back by last_census (synchronizes the stream).
  void set_census(bool on) { census_on_ = on; }
  void last_census(u64 out[4]);

 private:
  // k_color_reads (bitmap + count per read) and, with an include flag, the
  // quirk-2 pass over reads that found nothing
  void launch_threshold(const u64 *d_results, const u64 *d_kmer_ends, u64 n_reads,
                        const ColorParams &p, u32 *d_stats);
  const DeviceColors *colors_;
  hipStream_t stream_;
  bool own_stream_;
  u64 max_indexes_, max_reads_;
  u64 *d_bitmap_ = nullptr;   // [max_reads][words per read]
  u64 *d_nout_ = nullptr;     // per read
  u64 *d_scan_tmp_ = nullptr;
  u64 *d_quirk_ = nullptr;    // [4][max_reads]: blank_min, colored, colored_incl, slot_read
  u32 *d_global_cnt_ = nullptr;  // per-wave counters when the colors do not fit LDS
  u32 item_kmers_ = 0;
  int input_ = 0;             // key nodes, masks, ranks or their chains (color_search.hip)
  const DeviceIndex *chain_index_ = nullptr;  // the index of a key-chain input
  u64 *d_items_ = nullptr;     // [4][max_reads]: items, item ends, multi flags, their scan
  u32 *d_long_cnt_ = nullptr;  // [slots][num_colors] partial counts of multi-item reads
  u32 *d_long_meta_ = nullptr; // [slots][4]: finished items, found, not found, invalid
  u64 long_slots_ = 0;
  // staging for the reference interface
  u64 *d_warped_ = nullptr, *d_intervals_ = nullptr, *d_counts_ = nullptr;
  u64 *d_census_ = nullptr;    // [4] census of the last call
  bool census_on_ = false;
  hipEvent_t ev_[2];
  bool recorded_ = false;
  bool timing_ = true;
};

}  // namespace sbwt_amd
