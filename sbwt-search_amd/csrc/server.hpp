// Serving loop: the index (and color index) stay resident; requests (batches
// of reads) are queued and admitted to worker slots as slots and device memory
// free up, so requests that arrive while others run overlap on the device.
// No counterpart in the reference, which processes files in fixed streams.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "color_search.hpp"
#include "index_search.hpp"

namespace sbwt_amd {

struct ServerConfig {
  u32 slots = 4;                    // requests in flight on the device at once
  // chars per batch; a longer read is searched in pieces of this size that
  // overlap by k - 1 chars (at least k chars per slot)
  u64 max_chars_per_slot = 1 << 25;
  u64 max_reads_per_slot = 1 << 20;
  double memory_fraction = 0.9;     // of free device memory the slots may take
  bool move_to_key_kmer = false;
  ColorParams color;
  // Batches of up to kGraphMaxChars chars replay a HIP graph captured once per
  // (chars, reads) power-of-two bucket and slot: copies in, search, color
  // stage and copies out in one launch.
  bool graphs = true;
  // Copy every k-mer's index result back to the request (8 bytes per k-mer);
  // off, a request with a color index returns only its color lists.
  bool kmer_results = true;
};

constexpr u64 kGraphMaxChars = 1 << 20;
constexpr u64 kGraphMinChars = 1 << 12, kGraphMinReads = 1 << 6;

struct Request {
  u64 id = 0;
  // the request's reads: in_chars[0, n_chars), read ends in_ends[i] - ends_base
  // (either the request's own copies below, or caller memory that outlives the
  // request, see submit_borrowed)
  const char *in_chars = nullptr;
  const u64 *in_ends = nullptr;
  u64 n_chars = 0, n_reads = 0, ends_base = 0;
  std::vector<char> chars;
  std::vector<u64> read_ends;
  // results
  std::vector<u64> kmer_results;
  std::vector<u64> color_offsets;
  std::vector<u32> color_ids;
  std::string error;
  bool done = false;
  double submit_s = 0, start_s = 0, end_s = 0;  // seconds since server start
};

class Server {
 public:
  Server(const DeviceIndex *index, const DeviceColors *colors, const ServerConfig &cfg);
  ~Server();
  u64 submit(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads);
  // The same without copying: chars and read_ends (cumulative, each minus
  // ends_base is the end within chars) stay valid until the request is released.
  u64 submit_borrowed(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads,
                      u64 ends_base);
  // The clock origin of submit_s / start_s / end_s.
  std::chrono::steady_clock::time_point clock_origin() const { return t0_; }
  // Load generator (measurement): request i = reads [first_read[i],
  // first_read[i + 1]) of one concatenated batch (chars, cumulative read_ends),
  // submitted without copies from this thread at arrival_s[i] seconds after the
  // call starts while another thread waits for the requests in order. Per
  // request: latency_s = end - scheduled arrival (time spent behind a late
  // submitter counts), lag_s = submit - scheduled arrival, digest (nullable) =
  // a 64-bit hash of its k-mer results and color lists. wall_s = the call's
  // duration. Throws the first failed request's error after all have ended.
  void replay(const char *chars, const u64 *read_ends, const u64 *first_read, u64 n_requests,
              const double *arrival_s, double *latency_s, double *lag_s, u64 *digest,
              double *wall_s);
  // Blocks until request id is done; the request stays until release().
  std::shared_ptr<Request> wait(u64 id);
  void release(u64 id);
  u32 slots() const { return static_cast<u32>(slots_.size()); }
  u64 admitted_peak() const { return peak_in_flight_; }

 private:
  struct Slot {
    std::unique_ptr<Searcher> searcher;
    std::unique_ptr<ColorSearcher> colors;
    hipStream_t stream = nullptr;
    char *h_chars = nullptr;   // pinned staging
    u64 *h_ends = nullptr;
    char *d_chars = nullptr;
    u64 *d_ends = nullptr;
    u64 *d_results = nullptr;
    u64 *d_offsets = nullptr;
    u32 *d_ids = nullptr;
    u64 ids_capacity = 0;
    u64 *h_results = nullptr;  // pinned: results, color offsets and ids come back here
    u64 *h_offsets = nullptr;
    u32 *h_ids = nullptr;
    u32 *d_piece_counts = nullptr;  // a long read's piece: color counts, {found, not found, invalid}
    u32 *d_piece_stats = nullptr;
    u32 *h_piece = nullptr;         // pinned: the piece's counts then its stats
    // by (chars, reads) bucket, search algorithm and work-item size
    std::map<std::tuple<u64, u64, int, u32>, hipGraphExec_t> graphs;
    std::thread worker;
  };
  void run_slot(Slot &s);
  void process(Slot &s, Request &r);
  // One batch of whole reads staged in s.h_chars / s.h_ends (nc chars, nr
  // reads, the request's reads from read0); its results go to the request at
  // k-mer kmer_out (and its color lists). Returns the batch's k-mers.
  u64 run_batch(Slot &s, Request &r, u64 nc, u64 nr, u64 read0, u64 kmer_out);
  // Read i, longer than a slot, in pieces; returns its k-mers.
  u64 run_long_read(Slot &s, Request &r, u64 i, u64 kmer_out);
  u64 run_batch_graph(Slot &s, Request &r, u64 nc, u64 nr, u64 read0, u64 kmer_out, u64 cb,
                      u64 rb);
  hipGraphExec_t graph_for(Slot &s, u64 cb, u64 rb);
  double now() const;
  u64 enqueue(std::shared_ptr<Request> r);

  const DeviceIndex *index_;
  const DeviceColors *colors_;
  ServerConfig cfg_;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<std::shared_ptr<Request>> queue_;
  std::map<u64, std::shared_ptr<Request>> requests_;
  u64 next_id_ = 1;
  std::atomic<u64> in_flight_{0};
  u64 peak_in_flight_ = 0;
  bool stop_ = false;
  std::chrono::steady_clock::time_point t0_;
};

}  // namespace sbwt_amd
