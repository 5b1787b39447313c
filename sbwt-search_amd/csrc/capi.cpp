// extern "C" boundary of libsbwt_amd.so (declared in include/sbwt_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>

#include "../../include/sbwt_amd.h"
#include "color_search.hpp"
#include "device_index.hpp"
#include "format.hpp"
#include "index_parse.hpp"
#include "index_search.hpp"
#include "loaders.hpp"
#include "probe.hpp"
#include "seq_parse.hpp"
#include "server.hpp"

using sbwt_amd::DeviceIndex;
using sbwt_amd::Searcher;

struct sbwt_index {
  DeviceIndex dev;
  ~sbwt_index() { dev.release(); }
};

struct sbwt_searcher {
  std::unique_ptr<Searcher> impl;
};

struct sbwt_color_index {
  sbwt_amd::DeviceColors dev;
  ~sbwt_color_index() { dev.release(); }
};

struct sbwt_server {
  std::unique_ptr<sbwt_amd::Server> impl;
};

struct sbwt_formatter {
  std::unique_ptr<sbwt_amd::Formatter> impl;
};

struct sbwt_color_searcher {
  std::unique_ptr<sbwt_amd::ColorSearcher> impl;
  const sbwt_color_index *colors = nullptr;
};

struct sbwt_index_file_parser {
  std::unique_ptr<sbwt_amd::IndexFileParser> impl;
};

struct sbwt_color_formatter {
  std::unique_ptr<sbwt_amd::ColorFormatter> impl;
};

struct sbwt_fastq_parser {
  std::unique_ptr<sbwt_amd::FastqParser> impl;
  uint64_t n_chars = 0, n_reads = 0;  // of the last parse
};

struct sbwt_fasta_parser {
  std::unique_ptr<sbwt_amd::FastaParser> impl;
  hipEvent_t parsed = nullptr;  // the chars and read ends of the last parse are written
  ~sbwt_fasta_parser() {
    if (parsed) (void)hipEventDestroy(parsed);
  }
};

// The colors command's device stage: parser, color search and printer on the
// parser's stream, with the color id CSR between them.
struct sbwt_color_file_searcher {
  const sbwt_color_index *colors = nullptr;
  std::unique_ptr<sbwt_amd::IndexFileParser> parser;
  std::unique_ptr<sbwt_amd::ColorSearcher> search;
  std::unique_ptr<sbwt_amd::ColorFormatter> printer;
  uint64_t *d_offsets = nullptr;
  uint32_t *d_ids = nullptr;
  uint64_t ids_capacity = 0;
  ~sbwt_color_file_searcher() {
    if (d_offsets) (void)hipFree(d_offsets);
    if (d_ids) (void)hipFree(d_ids);
  }
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F f) {
  try {
    f();
    g_last_error.clear();
    return 0;
  } catch (const std::bad_alloc &) {
    g_last_error = "out of host memory";
    return -2;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
}

void require(bool ok, const char *msg) {
  if (!ok) throw std::runtime_error(msg);
}

// --threshold must lie in [0, 1] (ColorSearchArgumentParser.cpp:202-211).
sbwt_amd::ColorParams to_params(const sbwt_color_params *p) {
  sbwt_amd::ColorParams q;
  if (p) {
    if (!(p->threshold >= 0.0 && p->threshold <= 1.0))
      throw std::runtime_error("threshold must be in [0, 1]");
    q.threshold = p->threshold;
    q.include_not_found = p->include_not_found != 0;
    q.include_invalid = p->include_invalid != 0;
    q.first_read_quirk = p->first_read_quirk != 0;
  }
  return q;
}

}  // namespace

extern "C" {

const char *sbwt_last_error(void) { return g_last_error.c_str(); }

int sbwt_random_line_probe(uint64_t working_set_bytes, uint32_t blocks_per_cu, uint32_t steps,
                           float *ms, uint64_t *lines, uint64_t *table_bytes) {
  return guarded([&] {
    require(ms && lines && steps > 0, "null argument");
    const sbwt_amd::ProbeResult r = sbwt_amd::random_line_probe(working_set_bytes, blocks_per_cu, steps);
    *ms = r.ms;
    *lines = r.lines;
    if (table_bytes) *table_bytes = r.table_bytes;
  });
}

int sbwt_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
  return guarded([&] {
    size_t f = 0, t = 0;
    SBWT_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  });
}

int sbwt_index_create(const uint64_t *const acgt[4], uint64_t num_bits, uint32_t kmer_size,
                      const uint64_t *key_kmer_marks, const uint64_t *suffix_group_starts,
                      uint32_t presearch_letters, sbwt_index **out) {
  return guarded([&] {
    require(out != nullptr && acgt != nullptr, "null argument");
    require(num_bits >= 1, "an SBWT has at least one node");
    require(kmer_size >= 1 && kmer_size <= 32, "k-mer size must be in [1, 32]");
    for (int c = 0; c < 4; ++c) require(acgt[c] != nullptr, "null bit vector");
    const uint32_t kp = presearch_letters;  // 0: chosen from the index size
    require(kp <= kmer_size && kp <= 16, "presearch letters must be in [1, min(k, 16)]");
    auto idx = std::make_unique<sbwt_index>();
    idx->dev.build(acgt, num_bits, kmer_size, kp, key_kmer_marks, key_kmer_marks ? num_bits : 0,
                   suffix_group_starts);
    (void)sbwt_amd::noalloc_scratch(idx->dev.device);  // resolve the device's scratch at set-up
    *out = idx.release();
  });
}

int sbwt_index_load(const char *sbwt_path, const char *colors_path, uint32_t presearch_letters,
                    sbwt_index **out) {
  return guarded([&] {
    require(out != nullptr && sbwt_path != nullptr, "null argument");
    sbwt_amd::HostSbwt s = sbwt_amd::load_sbwt(sbwt_path);
    sbwt_amd::HostBitVector km;
    if (colors_path) {
      km = sbwt_amd::load_key_kmer_marks(colors_path);
      require(km.bits == s.num_bits, "key-k-mer marks and SBWT differ in length");
    }
    const uint64_t *acgt[4] = {s.acgt[0].data(), s.acgt[1].data(), s.acgt[2].data(), s.acgt[3].data()};
    const uint32_t k = static_cast<uint32_t>(s.k);
    const uint32_t kp = presearch_letters;  // 0: chosen from the index size
    require(kp <= k && kp <= 16, "presearch letters must be in [1, min(k, 16)]");
    auto idx = std::make_unique<sbwt_index>();
    idx->dev.build(acgt, s.num_bits, k, kp, colors_path ? km.words.data() : nullptr,
                   colors_path ? km.bits : 0, s.suffix_group_starts.data());
    (void)sbwt_amd::noalloc_scratch(idx->dev.device);  // resolve the device's scratch at set-up
    *out = idx.release();
  });
}

int sbwt_index_search_reads_noalloc(const sbwt_index *index, int move_to_key_kmer,
                                    const char *d_chars, uint64_t n_chars,
                                    const uint64_t *d_read_ends, uint64_t n_reads,
                                    uint64_t *d_results, void *hip_stream) {
  return guarded([&] {
    require(index != nullptr, "null argument");
    require(n_reads == 0 || (d_read_ends && d_results && (n_chars == 0 || d_chars)),
            "null argument");
    sbwt_amd::search_reads_noalloc(index->dev, move_to_key_kmer != 0, d_chars, n_chars,
                                   d_read_ends, n_reads, d_results,
                                   static_cast<hipStream_t>(hip_stream));
  });
}

int sbwt_index_build_kmer_table(sbwt_index *index, int *built) {
  return guarded([&] {
    require(index, "null argument");
    const bool ok = index->dev.build_kmer_hash();
    if (built) *built = ok ? 1 : 0;
  });
}

void sbwt_index_destroy(sbwt_index *index) { delete index; }

int sbwt_index_get_info(const sbwt_index *index, sbwt_index_info *info) {
  return guarded([&] {
    require(index && info, "null argument");
    info->num_nodes = index->dev.n;
    info->kmer_size = index->dev.k;
    info->presearch_letters = index->dev.kp;
    for (int i = 0; i < 5; ++i) info->c_map[i] = index->dev.c_map[i];
    info->device_bytes = index->dev.device_bytes;
    info->has_key_kmer_marks = index->dev.has_marks;
    info->wide = index->dev.wide;
    info->has_suffix_groups = index->dev.has_groups;
    info->suffix_groups_rejected = index->dev.groups_rejected;
  });
}

int sbwt_suffix_groups_consistent(const uint64_t *const acgt[4], const uint64_t *suffix_group_starts,
                                  uint64_t num_bits) {
  int ok = 0;
  const int rc = guarded([&] {
    require(acgt != nullptr && suffix_group_starts != nullptr, "null argument");
    for (int c = 0; c < 4; ++c) require(acgt[c] != nullptr, "null bit vector");
    ok = sbwt_amd::suffix_groups_consistent(acgt, suffix_group_starts, num_bits) ? 1 : 0;
  });
  return rc < 0 ? rc : ok;
}

int sbwt_searcher_create(sbwt_index *index, uint64_t max_chars_per_batch,
                         uint64_t max_reads_per_batch, int move_to_key_kmer, void *hip_stream,
                         sbwt_searcher **out) {
  return guarded([&] {
    require(index && out, "null argument");
    require(max_chars_per_batch > 0 && max_reads_per_batch > 0, "Not enough memory");
    auto s = std::make_unique<sbwt_searcher>();
    s->impl = std::make_unique<Searcher>(&index->dev, max_chars_per_batch, max_reads_per_batch,
                                         move_to_key_kmer != 0, static_cast<hipStream_t>(hip_stream));
    *out = s.release();
  });
}

void sbwt_searcher_destroy(sbwt_searcher *searcher) { delete searcher; }

int sbwt_searcher_sync(sbwt_searcher *searcher) {
  return guarded([&] {
    require(searcher, "null argument");
    SBWT_HIP(hipStreamSynchronize(searcher->impl->stream()));
  });
}

int sbwt_searcher_search(sbwt_searcher *searcher, const uint64_t *bit_seqs, uint64_t n_words,
                         const uint64_t *kmer_positions, uint64_t n_positions, uint64_t *results) {
  return guarded([&] {
    require(searcher && (n_positions == 0 || (bit_seqs && kmer_positions && results)), "null argument");
    searcher->impl->search_positions_host(bit_seqs, n_words, kmer_positions, n_positions, results);
  });
}

int sbwt_searcher_search_device(sbwt_searcher *searcher, const uint64_t *d_bit_seqs,
                                const uint64_t *d_kmer_positions, uint64_t n_positions,
                                uint64_t *d_results) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->search_positions_device(d_bit_seqs, d_kmer_positions, n_positions, d_results);
  });
}

int sbwt_searcher_search_reads(sbwt_searcher *searcher, const char *chars, uint64_t n_chars,
                               const uint64_t *read_ends, uint64_t n_reads, uint64_t *results,
                               uint64_t *n_kmers) {
  return guarded([&] {
    require(searcher && (n_reads == 0 || (read_ends && results)), "null argument");
    require(n_chars == 0 || chars, "null argument");
    require(n_reads == 0 || read_ends[n_reads - 1] == n_chars, "read_ends must end at n_chars");
    const uint64_t nk = searcher->impl->search_reads_host(chars, n_chars, read_ends, n_reads, results);
    if (n_kmers) *n_kmers = nk;
  });
}

int sbwt_searcher_search_reads_device(sbwt_searcher *searcher, const char *d_chars,
                                      uint64_t n_chars, const uint64_t *d_read_ends,
                                      uint64_t n_reads, uint64_t *d_results, uint64_t *n_kmers) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->search_reads_device(d_chars, n_chars, d_read_ends, n_reads, d_results, n_kmers);
  });
}

int sbwt_searcher_last_key_runs_time(sbwt_searcher *searcher, float *ms) {
  return guarded([&] {
    require(searcher && ms, "null argument");
    *ms = searcher->impl->last_timings().key_runs_ms;
  });
}

int sbwt_searcher_last_timings(sbwt_searcher *searcher, float *pack_ms, float *offsets_ms,
                               float *search_ms) {
  return guarded([&] {
    require(searcher, "null argument");
    const sbwt_amd::SearchTimings t = searcher->impl->last_timings();
    if (pack_ms) *pack_ms = t.pack_ms;
    if (offsets_ms) *offsets_ms = t.offsets_ms;
    if (search_ms) *search_ms = t.search_ms;
  });
}

int sbwt_searcher_set_stats(sbwt_searcher *searcher, int enable) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_stats(enable != 0);
  });
}

int sbwt_searcher_set_algorithm(sbwt_searcher *searcher, int algorithm) {
  return guarded([&] {
    require(searcher, "null argument");
    require(algorithm == SBWT_ALGO_STREAMING || algorithm == SBWT_ALGO_PER_KMER ||
                algorithm == SBWT_ALGO_AUTO || algorithm == SBWT_ALGO_STREAMING_ADAPTIVE,
            "unknown algorithm");
    searcher->impl->set_algorithm(algorithm);
  });
}

int sbwt_searcher_set_launch(sbwt_searcher *searcher, uint32_t chunk_kmers, uint32_t blocks_per_cu) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_launch(chunk_kmers, blocks_per_cu);
  });
}

int sbwt_searcher_get_launch(sbwt_searcher *searcher, uint32_t *chunk_kmers, uint32_t *blocks_per_cu) {
  return guarded([&] {
    require(searcher && chunk_kmers && blocks_per_cu, "null argument");
    const sbwt_amd::LaunchShape l = searcher->impl->launch();
    *chunk_kmers = l.chunk_kmers;
    *blocks_per_cu = l.blocks_per_cu;
  });
}

int sbwt_searcher_autotune(sbwt_searcher *searcher, const char *d_chars, uint64_t n_chars,
                           const uint64_t *d_read_ends, uint64_t n_reads, uint64_t *d_results,
                           uint32_t *chunk_kmers, uint32_t *blocks_per_cu, float *search_ms) {
  return guarded([&] {
    require(searcher && d_read_ends && d_results && (d_chars || n_chars == 0), "null argument");
    require(n_reads > 0, "autotune needs a batch with at least one read");
    const sbwt_amd::LaunchShape l =
        searcher->impl->autotune(d_chars, n_chars, d_read_ends, n_reads, d_results);
    if (chunk_kmers) *chunk_kmers = l.chunk_kmers;
    if (blocks_per_cu) *blocks_per_cu = l.blocks_per_cu;
    if (search_ms) *search_ms = l.search_ms;
  });
}

int sbwt_searcher_last_stats(sbwt_searcher *searcher, uint64_t stats[16]) {
  return guarded([&] {
    require(searcher && stats, "null argument");
    searcher->impl->last_stats(stats);
  });
}

int sbwt_pack_2bit_device(const char *d_chars, uint64_t n_chars, uint64_t *d_words,
                          uint32_t *d_invalid, void *hip_stream) {
  return guarded([&] {
    sbwt_amd::launch_pack(d_chars, n_chars, d_words, d_invalid, static_cast<hipStream_t>(hip_stream));
  });
}

int sbwt_host_alloc(uint64_t bytes, void **out) {
  return guarded([&] {
    require(out != nullptr, "null argument");
    *out = nullptr;
    SBWT_HIP(hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault));
  });
}

void sbwt_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int sbwt_formatter_create(uint64_t max_kmers, uint64_t max_pieces, void *hip_stream,
                          sbwt_formatter **out) {
  return guarded([&] {
    require(out != nullptr, "null argument");
    require(max_pieces >= 1, "max_pieces must be at least 1");
    auto f = std::make_unique<sbwt_formatter>();
    f->impl = std::make_unique<sbwt_amd::Formatter>(max_kmers, max_pieces,
                                                    static_cast<hipStream_t>(hip_stream));
    *out = f.release();
  });
}

void sbwt_formatter_destroy(sbwt_formatter *formatter) { delete formatter; }

int sbwt_formatter_format(sbwt_formatter *formatter, int mode, const uint64_t *d_results,
                          const uint64_t *d_piece_kmer_ends, const uint8_t *d_piece_closes,
                          uint64_t n_pieces, int line_open, char *d_out, uint64_t capacity,
                          uint64_t *n_bytes) {
  return guarded([&] {
    require(formatter && n_bytes, "null argument");
    require(mode == SBWT_FORMAT_ASCII || mode == SBWT_FORMAT_BINARY || mode == SBWT_FORMAT_BOOL,
            "unknown output format");
    require(n_pieces == 0 || (d_piece_kmer_ends && d_piece_closes), "null argument");
    *n_bytes = formatter->impl->format(mode, d_results, d_piece_kmer_ends, d_piece_closes, n_pieces,
                                       line_open != 0, d_out, capacity);
  });
}

int sbwt_searcher_search_reads_format(sbwt_searcher *searcher, sbwt_formatter *formatter, int mode,
                                      const char *chars, uint64_t n_chars,
                                      const uint64_t *read_ends, const uint8_t *read_closes,
                                      uint64_t n_reads, int line_open, char *out,
                                      uint64_t capacity, uint64_t *n_bytes) {
  return guarded([&] {
    require(searcher && formatter && n_bytes, "null argument");
    require(mode == SBWT_FORMAT_ASCII || mode == SBWT_FORMAT_BINARY || mode == SBWT_FORMAT_BOOL,
            "unknown output format");
    require(n_reads == 0 || (read_ends && read_closes && (chars || n_chars == 0)), "null argument");
    *n_bytes = 0;
    if (n_reads == 0) return;
    searcher->impl->search_reads_host(chars, n_chars, read_ends, n_reads, nullptr);
    *n_bytes = formatter->impl->format_to_host(mode, searcher->impl->results_device(),
                                               searcher->impl->kmer_ends_device(), read_closes,
                                               n_reads, line_open != 0, out, capacity);
  });
}

int sbwt_formatter_copy_last(sbwt_formatter *formatter, char *out, uint64_t capacity) {
  return guarded([&] {
    require(formatter && (out || capacity == 0), "null argument");
    formatter->impl->copy_last(out, capacity);
  });
}

int sbwt_searcher_kmer_ends_device(sbwt_searcher *searcher, const uint64_t **d_kmer_ends) {
  return guarded([&] {
    require(searcher && d_kmer_ends, "null argument");
    *d_kmer_ends = searcher->impl->kmer_ends_device();
  });
}

int sbwt_color_index_load(const char *tcolors_path, sbwt_color_index **out) {
  return guarded([&] {
    require(tcolors_path && out, "null argument");
    sbwt_amd::HostColors host = sbwt_amd::load_colors(tcolors_path);
    auto ci = std::make_unique<sbwt_color_index>();
    ci->dev.build(host);
    *out = ci.release();
  });
}

void sbwt_color_index_destroy(sbwt_color_index *colors) { delete colors; }

int sbwt_color_index_get_info(const sbwt_color_index *colors, sbwt_color_index_info *info) {
  return guarded([&] {
    require(colors && info, "null argument");
    info->num_colors = colors->dev.num_colors;
    info->num_color_sets = colors->dev.num_color_sets;
    info->num_nodes = colors->dev.num_nodes;
    info->device_bytes = colors->dev.device_bytes;
  });
}

int sbwt_color_searcher_create(sbwt_color_index *colors, uint64_t max_indexes_per_batch,
                               uint64_t max_reads_per_batch, void *hip_stream,
                               sbwt_color_searcher **out) {
  return guarded([&] {
    require(colors && out, "null argument");
    require(max_reads_per_batch > 0, "Not enough memory");
    auto s = std::make_unique<sbwt_color_searcher>();
    s->colors = colors;
    s->impl = std::make_unique<sbwt_amd::ColorSearcher>(&colors->dev, max_indexes_per_batch,
                                                        max_reads_per_batch,
                                                        static_cast<hipStream_t>(hip_stream));
    *out = s.release();
  });
}

void sbwt_color_searcher_destroy(sbwt_color_searcher *searcher) { delete searcher; }

int sbwt_color_searcher_search(sbwt_color_searcher *searcher, const uint64_t *warped_idxs,
                               uint64_t n_idxs, const uint64_t *warp_intervals,
                               uint64_t n_intervals, uint64_t *results) {
  return guarded([&] {
    require(searcher, "null argument");
    require(n_idxs == 0 || (warped_idxs && warp_intervals && results), "null argument");
    searcher->impl->search_warped_host(warped_idxs, n_idxs, warp_intervals, n_intervals, results);
  });
}

int sbwt_color_searcher_pseudoalign_device(sbwt_color_searcher *searcher,
                                           const uint64_t *d_kmer_results,
                                           const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                           const sbwt_color_params *params, uint64_t *d_offsets,
                                           uint32_t *d_ids, uint64_t ids_capacity,
                                           uint32_t *d_read_stats, uint64_t *n_ids) {
  return guarded([&] {
    require(searcher, "null argument");
    const uint64_t n = searcher->impl->pseudoalign_device(d_kmer_results, d_read_kmer_ends, n_reads,
                                                          to_params(params), d_offsets, d_ids,
                                                          ids_capacity, d_read_stats);
    if (n_ids) *n_ids = n;
  });
}

int sbwt_color_searcher_pseudoalign(sbwt_color_searcher *searcher, const uint64_t *kmer_results,
                                    const uint64_t *read_kmer_ends, uint64_t n_reads,
                                    const sbwt_color_params *params, uint64_t *offsets,
                                    uint32_t *ids, uint64_t ids_capacity, uint32_t *read_stats,
                                    uint64_t *n_ids) {
  return guarded([&] {
    require(searcher, "null argument");
    if (n_ids) *n_ids = 0;
    if (n_reads == 0) return;
    require(read_kmer_ends && offsets && (ids || ids_capacity == 0), "null argument");
    const uint64_t nk = read_kmer_ends[n_reads - 1];
    for (uint64_t r = 1; r < n_reads; ++r)
      require(read_kmer_ends[r] >= read_kmer_ends[r - 1], "read_kmer_ends must be non-decreasing");
    require(nk == 0 || kmer_results, "null argument");
    const sbwt_amd::ColorParams p = to_params(params);
    hipStream_t st = searcher->impl->stream();
    uint64_t *d_res = nullptr, *d_ends = nullptr, *d_off = nullptr;
    uint32_t *d_ids = nullptr, *d_stats = nullptr;
    auto cleanup = [&] {
      void *b[] = {d_res, d_ends, d_off, d_ids, d_stats};
      for (void *x : b)
        if (x) (void)hipFree(x);
    };
    try {
      SBWT_HIP(hipMalloc(&d_res, std::max<uint64_t>(nk, 1) * 8));
      SBWT_HIP(hipMalloc(&d_ends, n_reads * 8));
      SBWT_HIP(hipMalloc(&d_off, (n_reads + 1) * 8));
      SBWT_HIP(hipMalloc(&d_ids, std::max<uint64_t>(ids_capacity, 1) * 4));
      SBWT_HIP(hipMalloc(&d_stats, n_reads * 3 * 4));
      if (nk) SBWT_HIP(hipMemcpyAsync(d_res, kmer_results, nk * 8, hipMemcpyHostToDevice, st));
      SBWT_HIP(hipMemcpyAsync(d_ends, read_kmer_ends, n_reads * 8, hipMemcpyHostToDevice, st));
      const uint64_t n = searcher->impl->pseudoalign_device(d_res, d_ends, n_reads, p, d_off, d_ids,
                                                            ids_capacity, d_stats);
      SBWT_HIP(hipMemcpyAsync(offsets, d_off, (n_reads + 1) * 8, hipMemcpyDeviceToHost, st));
      if (n) SBWT_HIP(hipMemcpyAsync(ids, d_ids, n * 4, hipMemcpyDeviceToHost, st));
      if (read_stats)
        SBWT_HIP(hipMemcpyAsync(read_stats, d_stats, n_reads * 3 * 4, hipMemcpyDeviceToHost, st));
      SBWT_HIP(hipStreamSynchronize(st));
      if (n_ids) *n_ids = n;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int sbwt_index_build_color_table(sbwt_index *index, const sbwt_color_index *colors, int *kind) {
  return guarded([&] {
    require(index && colors, "null argument");
    sbwt_amd::build_color_table(index->dev, colors->dev);
    if (kind) *kind = static_cast<int>(index->dev.color_table_kind);
  });
}

int sbwt_searcher_set_color_output(sbwt_searcher *searcher, int on) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_color_output(on);
  });
}

int sbwt_color_searcher_set_input(sbwt_color_searcher *searcher, const sbwt_index *index) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_input(index ? &index->dev : nullptr);
  });
}

int sbwt_color_searcher_set_chain_input(sbwt_color_searcher *searcher, const sbwt_index *index) {
  return guarded([&] {
    require(searcher && index, "null argument");
    searcher->impl->set_input(&index->dev, true);
  });
}

int sbwt_color_searcher_autotune(sbwt_color_searcher *searcher, const uint64_t *d_kmer_results,
                                 const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                 const sbwt_color_params *params, uint64_t *d_offsets,
                                 uint32_t *d_ids, uint64_t ids_capacity, uint32_t *item_kmers,
                                 float *ms) {
  return guarded([&] {
    require(searcher, "null argument");
    const uint32_t ik = searcher->impl->autotune_items(d_kmer_results, d_read_kmer_ends, n_reads,
                                                       to_params(params), d_offsets, d_ids,
                                                       ids_capacity, ms);
    if (item_kmers) *item_kmers = ik;
  });
}

int sbwt_color_searcher_set_work_items(sbwt_color_searcher *searcher, uint32_t item_kmers) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_work_items(item_kmers);
  });
}

int sbwt_color_searcher_counts_device(sbwt_color_searcher *searcher,
                                      const uint64_t *d_kmer_results,
                                      const uint64_t *d_read_kmer_ends, uint64_t n_reads,
                                      uint32_t *d_counts) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->counts_device(d_kmer_results, d_read_kmer_ends, n_reads, d_counts);
  });
}

int sbwt_color_searcher_last_time(sbwt_color_searcher *searcher, float *ms) {
  return guarded([&] {
    require(searcher && ms, "null argument");
    *ms = searcher->impl->last_search_ms();
  });
}

int sbwt_color_searcher_set_census(sbwt_color_searcher *searcher, int enable) {
  return guarded([&] {
    require(searcher, "null argument");
    searcher->impl->set_census(enable != 0);
  });
}

int sbwt_color_searcher_last_census(sbwt_color_searcher *searcher, uint64_t census[4]) {
  return guarded([&] {
    require(searcher && census, "null argument");
    searcher->impl->last_census(census);
  });
}

int sbwt_index_file_parser_create(uint64_t max_bytes_per_batch, uint64_t max_reads_per_batch,
                                  void *hip_stream, sbwt_index_file_parser **out) {
  return guarded([&] {
    require(out, "null argument");
    require(max_reads_per_batch > 0, "max_reads_per_batch must be positive");
    auto p = std::make_unique<sbwt_index_file_parser>();
    p->impl = std::make_unique<sbwt_amd::IndexFileParser>(max_bytes_per_batch, max_reads_per_batch,
                                                          static_cast<hipStream_t>(hip_stream));
    *out = p.release();
  });
}

void sbwt_index_file_parser_destroy(sbwt_index_file_parser *parser) { delete parser; }

int sbwt_index_file_parser_parse_device(sbwt_index_file_parser *parser, int format,
                                        const char *d_body, uint64_t n_bytes,
                                        const uint64_t **d_results,
                                        const uint64_t **d_read_kmer_ends, uint64_t *n_results,
                                        uint64_t *n_reads) {
  return guarded([&] {
    require(parser && d_results && d_read_kmer_ends && n_results && n_reads, "null argument");
    require(n_bytes == 0 || d_body, "null argument");
    *n_reads = parser->impl->parse_device(format, d_body, n_bytes, n_results);
    *d_results = parser->impl->results();
    *d_read_kmer_ends = parser->impl->kmer_ends();
  });
}

int sbwt_index_file_parser_parse(sbwt_index_file_parser *parser, int format, const char *body,
                                 uint64_t n_bytes, uint64_t *results, uint64_t results_capacity,
                                 uint64_t *read_kmer_ends, uint64_t reads_capacity,
                                 uint64_t *n_results, uint64_t *n_reads) {
  return guarded([&] {
    require(parser && n_results && n_reads, "null argument");
    require(n_bytes == 0 || body, "null argument");
    uint64_t nr = 0;
    const uint64_t nl = parser->impl->parse_host(format, body, n_bytes, &nr);
    require(nr <= results_capacity && nl <= reads_capacity, "output arrays too small for the batch");
    require((nr == 0 || results) && (nl == 0 || read_kmer_ends), "null argument");
    hipStream_t st = parser->impl->stream();
    if (nr) SBWT_HIP(hipMemcpyAsync(results, parser->impl->results(), nr * 8, hipMemcpyDeviceToHost, st));
    if (nl)
      SBWT_HIP(hipMemcpyAsync(read_kmer_ends, parser->impl->kmer_ends(), nl * 8,
                              hipMemcpyDeviceToHost, st));
    SBWT_HIP(hipStreamSynchronize(st));
    *n_results = nr;
    *n_reads = nl;
  });
}

uint64_t sbwt_index_file_batch_cut(int format, const char *body, uint64_t n_bytes,
                                   uint64_t max_reads) {
  if (!body || (format != SBWT_INDEX_FILE_ASCII && format != SBWT_INDEX_FILE_BINARY)) return 0;
  return sbwt_amd::IndexFileParser::batch_cut(format, body, n_bytes, max_reads);
}

uint64_t sbwt_index_file_colored_cut(int format, const char *body, uint64_t n_bytes) {
  if (!body || (format != SBWT_INDEX_FILE_ASCII && format != SBWT_INDEX_FILE_BINARY)) return 0;
  return sbwt_amd::IndexFileParser::colored_cut(format, body, n_bytes);
}

int sbwt_fasta_parser_create(uint64_t max_bytes_per_chunk, uint64_t max_reads_per_chunk,
                             void *hip_stream, sbwt_fasta_parser **out) {
  return guarded([&] {
    require(out, "null argument");
    auto p = std::make_unique<sbwt_fasta_parser>();
    p->impl = std::make_unique<sbwt_amd::FastaParser>(max_bytes_per_chunk, max_reads_per_chunk,
                                                      static_cast<hipStream_t>(hip_stream));
    SBWT_HIP(hipEventCreateWithFlags(&p->parsed, hipEventDisableTiming));
    *out = p.release();
  });
}

void sbwt_fasta_parser_destroy(sbwt_fasta_parser *parser) { delete parser; }

int sbwt_fasta_parser_parse(sbwt_fasta_parser *parser, const char *raw, uint64_t n_bytes,
                            char *chars, uint64_t chars_capacity, uint64_t *read_ends,
                            uint64_t reads_capacity, uint64_t *n_chars, uint64_t *n_reads) {
  return guarded([&] {
    require(parser && n_chars && n_reads, "null argument");
    require(n_bytes == 0 || raw, "null argument");
    uint64_t nc = 0;
    const uint64_t nr = parser->impl->parse_host(raw, n_bytes, &nc);
    require(nc <= chars_capacity && nr <= reads_capacity, "output arrays too small for the chunk");
    require((nc == 0 || chars) && (nr == 0 || read_ends), "null argument");
    hipStream_t st = parser->impl->stream();
    if (nc) SBWT_HIP(hipMemcpyAsync(chars, parser->impl->chars(), nc, hipMemcpyDeviceToHost, st));
    if (nr)
      SBWT_HIP(hipMemcpyAsync(read_ends, parser->impl->read_ends(), nr * 8, hipMemcpyDeviceToHost, st));
    SBWT_HIP(hipStreamSynchronize(st));
    *n_chars = nc;
    *n_reads = nr;
  });
}

uint64_t sbwt_fasta_chunk_cut(const char *raw, uint64_t n_bytes, int last, uint64_t max_reads) {
  if (!raw) return 0;
  return sbwt_amd::FastaParser::chunk_cut(raw, n_bytes, last != 0, max_reads);
}

int sbwt_fasta_has_quality_line(const char *raw, uint64_t n_bytes, int at_line_start) {
  if (!raw) return 0;
  return sbwt_amd::FastaParser::has_quality_line(raw, n_bytes, at_line_start != 0) ? 1 : 0;
}

int sbwt_searcher_search_fasta_format(sbwt_searcher *searcher, sbwt_fasta_parser *parser,
                                      sbwt_formatter *formatter, int mode, const char *raw,
                                      uint64_t n_raw, char *out, uint64_t capacity,
                                      uint64_t *n_bytes, uint64_t *n_reads) {
  return guarded([&] {
    require(searcher && parser && formatter && n_bytes && n_reads, "null argument");
    require(mode == SBWT_FORMAT_ASCII || mode == SBWT_FORMAT_BINARY || mode == SBWT_FORMAT_BOOL,
            "unknown output format");
    require(n_raw == 0 || raw, "null argument");
    *n_bytes = 0;
    *n_reads = 0;
    uint64_t nc = 0;
    const uint64_t nr = parser->impl->parse_host(raw, n_raw, &nc);
    if (nr == 0) return;
    Searcher &s = *searcher->impl;
    SBWT_HIP(hipEventRecord(parser->parsed, parser->impl->stream()));
    SBWT_HIP(hipStreamWaitEvent(s.stream(), parser->parsed, 0));
    s.search_reads_device_staged(parser->impl->chars(), nc, parser->impl->read_ends(), nr);
    SBWT_HIP(hipStreamSynchronize(s.stream()));
    *n_bytes = formatter->impl->format_device_closes_to_host(
        mode, s.results_device(), s.kmer_ends_device(), parser->impl->ones(), nr, false, out,
        capacity);
    *n_reads = nr;
  });
}

int sbwt_fastq_parser_create(uint64_t max_bytes_per_chunk, uint64_t max_reads_per_chunk,
                             void *hip_stream, sbwt_fastq_parser **out) {
  return guarded([&] {
    require(out, "null argument");
    auto p = std::make_unique<sbwt_fastq_parser>();
    p->impl = std::make_unique<sbwt_amd::FastqParser>(max_bytes_per_chunk, max_reads_per_chunk,
                                                      static_cast<hipStream_t>(hip_stream));
    *out = p.release();
  });
}

void sbwt_fastq_parser_destroy(sbwt_fastq_parser *parser) { delete parser; }

int sbwt_fastq_parser_parse(sbwt_fastq_parser *parser, const char *raw, uint64_t n_bytes, int last,
                            char *chars, uint64_t chars_capacity, uint64_t *read_ends,
                            uint64_t reads_capacity, uint64_t *n_used, uint64_t *n_chars,
                            uint64_t *n_reads) {
  return guarded([&] {
    require(parser && n_used && n_chars && n_reads, "null argument");
    require(n_bytes == 0 || raw, "null argument");
    parser->n_chars = parser->n_reads = 0;
    uint64_t nr = 0, nc = 0;
    const uint64_t used = parser->impl->parse_host(raw, n_bytes, last != 0, &nr, &nc);
    hipStream_t st = parser->impl->stream();
    if (chars && nc) {
      require(nc <= chars_capacity, "output arrays too small for the chunk");
      SBWT_HIP(hipMemcpyAsync(chars, parser->impl->chars(), nc, hipMemcpyDeviceToHost, st));
    }
    if (read_ends && nr) {
      require(nr <= reads_capacity, "output arrays too small for the chunk");
      SBWT_HIP(hipMemcpyAsync(read_ends, parser->impl->read_ends(), nr * 8, hipMemcpyDeviceToHost, st));
    }
    SBWT_HIP(hipStreamSynchronize(st));
    parser->n_chars = nc;
    parser->n_reads = nr;
    *n_used = used;
    *n_chars = nc;
    *n_reads = nr;
  });
}

int sbwt_searcher_search_fastq_format(sbwt_searcher *searcher, sbwt_fastq_parser *parser,
                                      sbwt_formatter *formatter, int mode, char *out,
                                      uint64_t capacity, uint64_t *n_bytes) {
  return guarded([&] {
    require(searcher && parser && formatter && n_bytes, "null argument");
    require(mode == SBWT_FORMAT_ASCII || mode == SBWT_FORMAT_BINARY || mode == SBWT_FORMAT_BOOL,
            "unknown output format");
    *n_bytes = 0;
    if (parser->n_reads == 0) return;
    Searcher &s = *searcher->impl;
    // the parse ended with a synchronization of the parser's stream
    s.search_reads_device_staged(parser->impl->chars(), parser->n_chars,
                                 parser->impl->read_ends(), parser->n_reads);
    SBWT_HIP(hipStreamSynchronize(s.stream()));
    *n_bytes = formatter->impl->format_device_closes_to_host(
        mode, s.results_device(), s.kmer_ends_device(), parser->impl->ones(), parser->n_reads,
        false, out, capacity);
  });
}

int sbwt_color_formatter_create(uint64_t max_ids, uint64_t max_reads, void *hip_stream,
                                sbwt_color_formatter **out) {
  return guarded([&] {
    require(out, "null argument");
    auto f = std::make_unique<sbwt_color_formatter>();
    f->impl = std::make_unique<sbwt_amd::ColorFormatter>(max_ids, max_reads,
                                                         static_cast<hipStream_t>(hip_stream));
    *out = f.release();
  });
}

void sbwt_color_formatter_destroy(sbwt_color_formatter *formatter) { delete formatter; }

int sbwt_color_formatter_format(sbwt_color_formatter *formatter, int mode,
                                const uint64_t *d_offsets, const uint32_t *d_ids,
                                uint64_t n_reads, uint64_t num_colors, char *d_out,
                                uint64_t capacity, uint64_t *n_bytes) {
  return guarded([&] {
    require(formatter && n_bytes, "null argument");
    require(n_reads == 0 || (d_offsets && d_ids), "null argument");
    *n_bytes = formatter->impl->format(mode, d_offsets, d_ids, n_reads, num_colors, d_out, capacity);
  });
}

int sbwt_color_file_searcher_create(sbwt_color_index *colors, uint64_t max_bytes_per_batch,
                                    uint64_t max_reads_per_batch, void *hip_stream,
                                    sbwt_color_file_searcher **out) {
  return guarded([&] {
    require(colors && out, "null argument");
    require(max_reads_per_batch > 0, "max_reads_per_batch must be positive");
    auto f = std::make_unique<sbwt_color_file_searcher>();
    f->colors = colors;
    f->parser = std::make_unique<sbwt_amd::IndexFileParser>(max_bytes_per_batch, max_reads_per_batch,
                                                            static_cast<hipStream_t>(hip_stream));
    hipStream_t st = f->parser->stream();
    f->search = std::make_unique<sbwt_amd::ColorSearcher>(&colors->dev, f->parser->max_results(),
                                                          max_reads_per_batch, st);
    f->search->set_timing(false);
    // a read reports at most min(num_colors, its found results) ids
    const uint64_t nc = std::max<uint64_t>(colors->dev.num_colors, 1);
    f->ids_capacity = max_reads_per_batch > f->parser->max_results() / nc
                          ? f->parser->max_results()
                          : max_reads_per_batch * nc;
    f->printer = std::make_unique<sbwt_amd::ColorFormatter>(f->ids_capacity, max_reads_per_batch, st);
    SBWT_HIP(hipMalloc(&f->d_offsets, (max_reads_per_batch + 1) * 8));
    SBWT_HIP(hipMalloc(&f->d_ids, std::max<uint64_t>(f->ids_capacity, 1) * 4));
    *out = f.release();
  });
}

void sbwt_color_file_searcher_destroy(sbwt_color_file_searcher *searcher) { delete searcher; }

int sbwt_color_file_searcher_batch(sbwt_color_file_searcher *searcher, int in_format,
                                   const char *body, uint64_t n_bytes,
                                   const sbwt_color_params *params, int out_mode, char *out,
                                   uint64_t capacity, uint64_t *n_bytes_out, uint64_t *n_reads) {
  return guarded([&] {
    require(searcher && n_bytes_out && n_reads, "null argument");
    require(n_bytes == 0 || body, "null argument");
    require(out_mode == SBWT_COLOR_FORMAT_ASCII || out_mode == SBWT_COLOR_FORMAT_BINARY ||
                out_mode == SBWT_COLOR_FORMAT_CSV,
            "unknown color output format");
    const sbwt_amd::ColorParams p = to_params(params);
    *n_bytes_out = 0;
    uint64_t nr = 0;
    const uint64_t nl = searcher->parser->parse_host(in_format, body, n_bytes, &nr);
    *n_reads = nl;
    if (nl == 0) return;
    searcher->search->pseudoalign_device_async(searcher->parser->results(),
                                               searcher->parser->kmer_ends(), nl, p,
                                               searcher->d_offsets, searcher->d_ids,
                                               searcher->ids_capacity, nullptr);
    *n_bytes_out = searcher->printer->format_to_host(out_mode, searcher->d_offsets, searcher->d_ids,
                                                     nl, searcher->colors->dev.num_colors, out,
                                                     capacity);
  });
}

int sbwt_color_file_searcher_copy_last(sbwt_color_file_searcher *searcher, char *out,
                                       uint64_t capacity) {
  return guarded([&] {
    require(searcher && out, "null argument");
    searcher->printer->copy_last(out, capacity);
  });
}

int sbwt_server_create(sbwt_index *index, sbwt_color_index *colors,
                       const sbwt_server_config *config, sbwt_server **out) {
  return guarded([&] {
    require(index && config && out, "null argument");
    sbwt_amd::ServerConfig c;
    c.slots = config->slots;
    c.max_chars_per_slot = config->max_chars_per_slot;
    c.max_reads_per_slot = config->max_reads_per_slot;
    c.memory_fraction = config->memory_fraction;
    c.move_to_key_kmer = config->move_to_key_kmer != 0;
    c.graphs = config->no_graphs == 0;
    c.kmer_results = config->no_kmer_results == 0 || colors == nullptr;
    if (colors) c.color = to_params(&config->color);
    auto s = std::make_unique<sbwt_server>();
    s->impl = std::make_unique<sbwt_amd::Server>(&index->dev, colors ? &colors->dev : nullptr, c);
    *out = s.release();
  });
}

int sbwt_server_submit(sbwt_server *server, const char *chars, uint64_t n_chars,
                       const uint64_t *read_ends, uint64_t n_reads, uint64_t *request_id) {
  return guarded([&] {
    require(server && request_id && (n_chars == 0 || chars) && (n_reads == 0 || read_ends),
            "null argument");
    *request_id = server->impl->submit(chars, n_chars, read_ends, n_reads);
  });
}

int sbwt_server_wait(sbwt_server *server, uint64_t request_id, sbwt_server_result *out) {
  return guarded([&] {
    require(server && out, "null argument");
    auto r = server->impl->wait(request_id);
    out->n_kmers = r->kmer_results.size();
    out->kmer_results = r->kmer_results.data();
    out->n_reads = r->n_reads;
    out->color_offsets = r->color_offsets.empty() ? nullptr : r->color_offsets.data();
    out->color_ids = r->color_ids.data();
    out->submit_s = r->submit_s;
    out->start_s = r->start_s;
    out->end_s = r->end_s;
    if (!r->error.empty()) throw std::runtime_error(r->error);
  });
}

int sbwt_server_replay(sbwt_server *server, const char *chars, const uint64_t *read_ends,
                       const uint64_t *first_read, uint64_t n_requests, const double *arrival_s,
                       double *latency_s, double *lag_s, uint64_t *digest, double *wall_s) {
  return guarded([&] {
    require(server && first_read && arrival_s && latency_s && lag_s && wall_s, "null argument");
    require(first_read[0] == 0, "first_read must start at 0");
    for (uint64_t i = 0; i < n_requests; ++i) {
      require(first_read[i + 1] >= first_read[i], "first_read must be non-decreasing");
      require(i == 0 || arrival_s[i] >= arrival_s[i - 1], "arrival times must be non-decreasing");
    }
    require(first_read[n_requests] == 0 || (chars && read_ends), "null argument");
    server->impl->replay(chars, read_ends, first_read, n_requests, arrival_s, latency_s, lag_s,
                         digest, wall_s);
  });
}

int sbwt_server_release(sbwt_server *server, uint64_t request_id) {
  return guarded([&] {
    require(server, "null argument");
    server->impl->release(request_id);
  });
}

int sbwt_server_get_slots(sbwt_server *server, uint32_t *slots, uint64_t *peak_in_flight) {
  return guarded([&] {
    require(server, "null argument");
    if (slots) *slots = server->impl->slots();
    if (peak_in_flight) *peak_in_flight = server->impl->admitted_peak();
  });
}

void sbwt_server_destroy(sbwt_server *server) { delete server; }

}  // extern "C"
