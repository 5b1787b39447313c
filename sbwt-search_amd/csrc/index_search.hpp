// Index search: 2-bit packing, per-read k-mer offsets and the SBWT search
// kernels, driven by a Searcher that owns one HIP stream and its batch
// buffers (the reference's IndexSearcher, src/IndexSearcher/IndexSearcher.h:25-46).
#pragma once

#include "device_index.hpp"

namespace sbwt_amd {

// Search algorithms of the fused reads path: streaming persistent lanes
// (consecutive k-mers of a read step from the previous k-mer's node) or one
// independent search per k-mer (the reference's per-thread search).
constexpr int kAlgoStreaming = 0;
constexpr int kAlgoPerKmer = 1;
// Auto: per k-mer below kAutoStreamingChars chars, streaming above with work
// items shrunk until every persistent lane has one: chunk =
// clamp(ceil(chars / lanes), kAutoMinChunkKmers, the launch shape's chunk).
// Alone on the GPU, a batch of a few million chars finishes sooner per k-mer
// (every k-mer in flight at once, one search's chain of steps) although it
// reads one line per letter; above, streaming's few lines per k-mer win.
// Streaming-adaptive: the same shrunk work items from kAdaptiveStreamingChars
// chars up, for batches that share the GPU with others (a server under load),
// where the lines read, not one batch's chain of steps, bound the throughput.
constexpr int kAlgoAuto = 2;
constexpr int kAlgoStreamingAdaptive = 3;
constexpr u64 kAutoStreamingChars = 1ULL << 22;
// With the index's k-mer table a per-k-mer search reads about one bucket pair,
// so it stays ahead longer (measured on a 95 M-node index: 20.5 against 18.2 G
// k-mers/s at 15 M chars, 21.6 against 37.1 at 150 M).
constexpr u64 kAutoStreamingCharsTable = 1ULL << 24;
constexpr u64 kAdaptiveStreamingChars = 1ULL << 16;
constexpr u32 kAutoMinChunkKmers = 8;

struct SearchTimings {
  float pack_ms = 0, offsets_ms = 0, search_ms = 0;
  float key_runs_ms = 0;  // of search_ms: the key-k-mer runs after the streaming search
};

// Launch shape of the streaming search (work item = up to chunk_kmers
// consecutive k-mers of one read; blocks_per_cu persistent 256-lane
// workgroups per CU, each holding 16 KiB of LDS for staged results).
struct LaunchShape {
  u32 chunk_kmers;
  u32 blocks_per_cu;
  float search_ms;  // measured by autotune
};
constexpr u32 kMinChunkKmers = 64, kMaxChunkKmers = 65535, kDefaultChunkKmers = 256;
constexpr u32 kMaxBlocksPerCu = 8, kDefaultBlocksPerCu = 4;

class Searcher {
 public:
  Searcher(const DeviceIndex *index, u64 max_chars, u64 max_reads, bool move_to_key_kmer,
           hipStream_t stream);
  ~Searcher();

  // Reference interface: packed 2-bit sequence + explicit k-mer start
  // positions -> one u64 per position (node index or kNotFound), host
  // buffers, synchronous (IndexSearcher::search, IndexSearcher.cpp:26-51).
  void search_positions_host(const u64 *bit_seqs, u64 n_words, const u64 *positions, u64 n,
                             u64 *results);
  // Fused form: ASCII reads + cumulative read ends -> classified result per
  // k-mer (index / kNotFound / kInvalid), all buffers on the device, enqueued
  // on the searcher's stream. Returns the number of k-mers.
  u64 search_reads_device(const char *d_chars, u64 n_chars, const u64 *d_read_ends, u64 n_reads,
                          u64 *d_results, u64 *n_kmers_host);
  // Same with host buffers, synchronous.
  // results may be null: they then stay on the device (results_device()).
  u64 search_reads_host(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads,
                        u64 *results);
  // Device reads in, results kept in the searcher's own buffer
  // (results_device()); enqueued on the searcher's stream.
  void search_reads_device_staged(const char *d_chars, u64 n_chars, const u64 *d_read_ends,
                                  u64 n_reads);
  // Device positions form (explicit positions already on the device).
  void search_positions_device(const u64 *d_bit_seqs, const u64 *d_positions, u64 n,
                               u64 *d_results);
  // Launch shape of the streaming search: k-mers per work item and persistent
  // workgroups per CU. autotune times the search of one caller batch for each
  // shape of a small grid, keeps the fastest and returns it (results of the
  // batch are in d_results, the same for every shape).
  void set_launch(u32 chunk_kmers, u32 blocks_per_cu);
  LaunchShape launch() const { return {chunk_kmers_, blocks_per_cu_, 0.0f}; }
  // k-mers per work item of a streaming search over n_chars chars.
  u32 stream_chunk(u64 n_chars) const;
  // Adaptive work items assume 1 / share of the persistent lanes: a caller
  // running `share` batches at once on the GPU sizes each for its part.
  void set_lane_share(u32 share) { lane_share_ = share ? share : 1; }
  LaunchShape autotune(const char *d_chars, u64 n_chars, const u64 *d_read_ends, u64 n_reads,
                       u64 *d_results);

  hipStream_t stream() const { return stream_; }
  SearchTimings last_timings();
  void set_timing(bool on) { timing_ = on; }
  // Census mode: the fused search also counts its work per k-mer class (the
  // words are listed with sbwt_searcher_last_stats); last_stats reads them
  // (synchronous).
  void set_stats(bool on) { stats_ = on; }
  void set_algorithm(int a) { algorithm_ = a; }
  // Results of the reads path with the key-k-mer move: the key's node (the
  // reference's output) or, with color_output, the key's entry in the
  // index's color table (built from a color index), the input of the fused
  // color stage. Needs move_to_key_kmer and a built table.
  // kColorOutputNone (key nodes), kColorOutputEntries (each found k-mer's
  // color-table entry) or kColorOutputChains (the key chains, for a color
  // searcher with chain input: the key pass is left to it)
  void set_color_output(int mode);
  int color_output() const { return color_output_; }
  int algorithm() const { return algorithm_; }
  void last_stats(u64 out[kStatsWords]);
  const u64 *kmer_ends_device() const { return d_kmer_ends_; }
  const u64 *results_device() const { return d_results_; }
  u64 max_chars() const { return max_chars_; }
  u64 max_reads() const { return max_reads_; }

 private:
  void ensure_staging(u64 chars, u64 reads, u64 positions);
  const DeviceIndex *index_;
  hipStream_t stream_;
  bool own_stream_;
  bool move_to_key_;
  bool timing_ = true;
  bool stats_ = false;
  int algorithm_ = kAlgoAuto;
  int color_output_ = 0;
  IndexView reads_view() const;
  u32 lane_share_ = 1;
  u64 *d_stats_ = nullptr;
  u64 max_chunks_ = 0;
  ulonglong2 *d_chunks_ = nullptr;
  u64 *d_chunk_counts_ = nullptr;  // [unused | inclusive work-item ends per read]
  u64 *d_counter_ = nullptr;
  u32 persistent_blocks_ = 0;
  int cus_ = 1;
  u32 chunk_kmers_ = kDefaultChunkKmers;
  u32 blocks_per_cu_ = kDefaultBlocksPerCu;
  u64 max_chars_, max_reads_;
  u64 *d_packed_ = nullptr;
  u32 *d_invalid_ = nullptr;
  u64 *d_kmer_ends_ = nullptr;
  u64 *d_scan_tmp_ = nullptr;
  u64 scan_tmp_len_ = 0;
  // staging for host-buffer calls
  char *d_chars_ = nullptr;
  u64 *d_read_ends_ = nullptr;
  u64 *d_results_ = nullptr;
  u64 *d_positions_ = nullptr;
  u64 *d_bits_ = nullptr;
  u64 staged_chars_ = 0, staged_reads_ = 0, staged_positions_ = 0;
  hipEvent_t ev_[5];  // pack, offsets, search, end; [4] between the search and the key runs
  bool key_runs_timed_ = false;
  bool events_recorded_ = false;
};

// Device kernels exposed for reuse by the color stage and tests.
void launch_pack(const char *d_chars, u64 n_chars, u64 *d_packed, u32 *d_invalid, hipStream_t s);

// The fused per-k-mer search of one batch (the kernels of
// Searcher::search_reads_device with the per-k-mer algorithm) that allocates
// nothing: its packed sequence, invalid masks, k-mer ends and scan sums live
// in a file-scope device array of kNoAllocScratchBytes, one copy per GPU,
// whose address is resolved once per device (noalloc_scratch). Batches are
// limited to kNoAllocMaxChars chars and kNoAllocMaxReads reads. Every call on
// one GPU shares that GPU's array, so calls that overlap on a GPU race.
constexpr u64 kNoAllocMaxChars = 1ULL << 22;
constexpr u64 kNoAllocMaxReads = 1ULL << 18;
// Device address of this GPU's scratch array (the index's device must be
// current); resolved by hipGetSymbolAddress on the first call per device.
u64 *noalloc_scratch(int device);
void search_reads_noalloc(const DeviceIndex &index, bool move_to_key_kmer, const char *d_chars,
                          u64 n_chars, const u64 *d_read_ends, u64 n_reads, u64 *d_results,
                          hipStream_t stream);

}  // namespace sbwt_amd
