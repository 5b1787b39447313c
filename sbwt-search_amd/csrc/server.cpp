// Serving loop (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace sbwt_amd {

Server::Server(const DeviceIndex *index, const DeviceColors *colors, const ServerConfig &cfg)
    : index_(index), colors_(colors), cfg_(cfg), t0_(std::chrono::steady_clock::now()) {
  if (cfg_.slots == 0 || cfg_.max_chars_per_slot == 0 || cfg_.max_reads_per_slot == 0)
    throw std::runtime_error("Not enough memory");
  if (cfg_.move_to_key_kmer && !index_->has_marks)
    throw std::runtime_error("move to key k-mer requested but the index has no key-k-mer marks");
  if (cfg_.max_chars_per_slot < index_->k)
    throw std::runtime_error("a slot must hold at least k chars");
  const u64 nc = colors_ ? colors_->num_colors : 0;
  const u64 C = cfg_.max_chars_per_slot, R = cfg_.max_reads_per_slot;
  const u64 ids_cap = colors_ ? std::min<u64>(C, R * nc) : 0;
  // device bytes of one slot: inputs, results, searcher and color-searcher buffers
  const u64 per_slot = C * (1 + 8 + 1) + R * (8 * 6) + C / 256 * 16 +
                       (colors_ ? R * ((nc + 63) / 64 * 8 + 24) + ids_cap * 4 : 0) + (8u << 20);
  size_t free_b = 0, total = 0;
  SBWT_HIP(hipMemGetInfo(&free_b, &total));
  const u64 budget = static_cast<u64>(static_cast<double>(free_b) * cfg_.memory_fraction);
  const u64 fit = budget / per_slot;
  if (fit == 0) throw std::runtime_error("Not enough memory");
  const u32 n_slots = static_cast<u32>(std::min<u64>(cfg_.slots, fit));
  for (u32 i = 0; i < n_slots; ++i) {
    auto s = std::make_unique<Slot>();
    SBWT_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->searcher = std::make_unique<Searcher>(index_, C, R, cfg_.move_to_key_kmer, s->stream);
    s->searcher->set_timing(false);
    if (colors_) {
      s->colors = std::make_unique<ColorSearcher>(colors_, C, R, s->stream);
      s->colors->set_timing(false);
    }
    SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_chars), C + 64, hipHostMallocDefault));
    SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_ends), R * 8, hipHostMallocDefault));
    // pinned landing buffers for everything that comes back (copies from
    // device memory into pageable memory would go through a bounce buffer)
    if (cfg_.kmer_results)
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_results), C * 8 + 8, hipHostMallocDefault));
    SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_offsets), (R + 1) * 8, hipHostMallocDefault));
    if (colors_)
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_ids), std::max<u64>(ids_cap, 1) * 4,
                             hipHostMallocDefault));
    SBWT_HIP(hipMalloc(&s->d_chars, C + 64));
    SBWT_HIP(hipMalloc(&s->d_ends, R * 8));
    SBWT_HIP(hipMalloc(&s->d_results, (C + 1) * 8));
    if (colors_) {
      s->ids_capacity = std::max<u64>(ids_cap, 1);
      SBWT_HIP(hipMalloc(&s->d_offsets, (R + 1) * 8));
      SBWT_HIP(hipMalloc(&s->d_ids, s->ids_capacity * 4));
      SBWT_HIP(hipMalloc(&s->d_piece_counts, (nc + 1) * 4));
      SBWT_HIP(hipMalloc(&s->d_piece_stats, 16));
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_piece), (nc + 4) * 4, hipHostMallocDefault));
    }
    slots_.push_back(std::move(s));
  }
  for (auto &s : slots_) {
    Slot *sp = s.get();
    s->worker = std::thread([this, sp] { run_slot(*sp); });
  }
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto &s : slots_) {
    if (s->worker.joinable()) s->worker.join();
    for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.second);
    s->colors.reset();
    s->searcher.reset();
    void *dev[] = {s->d_chars, s->d_ends, s->d_results, s->d_offsets, s->d_ids,
                   s->d_piece_counts, s->d_piece_stats};
    for (void *p : dev)
      if (p) (void)hipFree(p);
    if (s->h_piece) (void)hipHostFree(s->h_piece);
    if (s->h_chars) (void)hipHostFree(s->h_chars);
    if (s->h_ends) (void)hipHostFree(s->h_ends);
    if (s->h_results) (void)hipHostFree(s->h_results);
    if (s->h_offsets) (void)hipHostFree(s->h_offsets);
    if (s->h_ids) (void)hipHostFree(s->h_ids);
    if (s->stream) (void)hipStreamDestroy(s->stream);
  }
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

u64 Server::submit(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads) {
  auto r = std::make_shared<Request>();
  r->chars.assign(chars, chars + n_chars);
  r->read_ends.assign(read_ends, read_ends + n_reads);
  r->in_chars = r->chars.data();
  r->in_ends = r->read_ends.data();
  r->n_chars = n_chars;
  r->n_reads = n_reads;
  return enqueue(std::move(r));
}

u64 Server::submit_borrowed(const char *chars, u64 n_chars, const u64 *read_ends, u64 n_reads,
                            u64 ends_base) {
  auto r = std::make_shared<Request>();
  r->in_chars = chars;
  r->in_ends = read_ends;
  r->n_chars = n_chars;
  r->n_reads = n_reads;
  r->ends_base = ends_base;
  return enqueue(std::move(r));
}

u64 Server::enqueue(std::shared_ptr<Request> r) {
  std::lock_guard<std::mutex> g(mu_);
  r->id = next_id_++;
  r->submit_s = now();
  requests_[r->id] = r;
  queue_.push_back(r);
  cv_work_.notify_one();
  return r->id;
}

std::shared_ptr<Request> Server::wait(u64 id) {
  std::unique_lock<std::mutex> g(mu_);
  auto it = requests_.find(id);
  if (it == requests_.end()) throw std::runtime_error("unknown request id " + std::to_string(id));
  auto r = it->second;
  cv_done_.wait(g, [&] { return r->done; });
  return r;
}

void Server::release(u64 id) {
  std::lock_guard<std::mutex> g(mu_);
  requests_.erase(id);
}

void Server::run_slot(Slot &s) {
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> g(mu_);
      cv_work_.wait(g, [&] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;  // stop_ and drained
      r = queue_.front();
      queue_.pop_front();
      peak_in_flight_ = std::max<u64>(peak_in_flight_, ++in_flight_);
    }
    r->start_s = now();
    try {
      process(s, *r);
    } catch (const std::exception &e) {
      r->error = e.what();
      (void)hipStreamSynchronize(s.stream);
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      r->end_s = now();
      r->done = true;
      --in_flight_;
    }
    cv_done_.notify_all();
  }
}

// One request through one slot: batches of whole reads up to the slot's
// capacity, each packed, searched (and color-searched) on the slot's stream.
void Server::process(Slot &s, Request &r) {
  const u64 n_reads = r.n_reads;
  const u64 n_chars = r.n_chars;
  const u64 k = index_->k;
  // read end i within the request's chars (validated below before any use)
  const auto end_of = [&](u64 i) { return r.in_ends[i] - r.ends_base; };
  u64 prev = 0, total_kmers = 0;
  for (u64 i = 0; i < n_reads; ++i) {
    if (r.in_ends[i] < r.ends_base || end_of(i) < prev || end_of(i) > n_chars)
      throw std::runtime_error("read_ends must be non-decreasing and end at n_chars");
    const u64 len = end_of(i) - prev;
    total_kmers += len >= k ? len - k + 1 : 0;
    prev = end_of(i);
  }
  if (n_reads && end_of(n_reads - 1) != n_chars) throw std::runtime_error("read_ends must end at n_chars");
  if (cfg_.kmer_results) r.kmer_results.resize(total_kmers);
  if (colors_) {
    r.color_offsets.assign(n_reads + 1, 0);
    r.color_ids.clear();
  }
  u64 read0 = 0, kmer_out = 0;
  while (read0 < n_reads) {
    // whole reads up to the slot capacity
    const u64 c0 = read0 ? end_of(read0 - 1) : 0;
    if (end_of(read0) - c0 > cfg_.max_chars_per_slot) {
      kmer_out += run_long_read(s, r, read0, kmer_out);
      ++read0;
      continue;
    }
    u64 read1 = read0;
    while (read1 < n_reads && read1 - read0 < cfg_.max_reads_per_slot &&
           end_of(read1) - c0 <= cfg_.max_chars_per_slot)
      ++read1;
    const u64 nr = read1 - read0, nc = end_of(read1 - 1) - c0;
    std::memcpy(s.h_chars, r.in_chars + c0, nc);
    for (u64 i = 0; i < nr; ++i) s.h_ends[i] = end_of(read0 + i) - c0;
    // power-of-two bucket of the batch: small batches replay a captured graph
    u64 cb = kGraphMinChars, rb = kGraphMinReads;
    while (cb < nc) cb <<= 1;
    while (rb < nr) rb <<= 1;
    const bool graph = cfg_.graphs && cb <= kGraphMaxChars && cb <= cfg_.max_chars_per_slot &&
                       rb <= cfg_.max_reads_per_slot && rb <= kGraphMaxChars;
    // alone on the GPU a batch takes the search of least latency; sharing it
    // with other requests, streaming (fewest lines read) with work items
    // sized for its share of the chip
    const u64 sharing = in_flight_.load();
    s.searcher->set_algorithm(sharing > 1 ? kAlgoStreamingAdaptive : kAlgoAuto);
    s.searcher->set_lane_share(static_cast<u32>(std::max<u64>(sharing, 1)));
    kmer_out += graph ? run_batch_graph(s, r, nc, nr, read0, kmer_out, cb, rb)
                      : run_batch(s, r, nc, nr, read0, kmer_out);
    read0 = read1;
  }
}

u64 Server::run_batch(Slot &s, Request &r, u64 nc, u64 nr, u64 read0, u64 kmer_out) {
  SBWT_HIP(hipMemcpyAsync(s.d_chars, s.h_chars, nc, hipMemcpyHostToDevice, s.stream));
  SBWT_HIP(hipMemcpyAsync(s.d_ends, s.h_ends, nr * 8, hipMemcpyHostToDevice, s.stream));
  u64 nk = 0;
  s.searcher->search_reads_device(s.d_chars, nc, s.d_ends, nr, s.d_results, &nk);
  if (nk && cfg_.kmer_results)
    SBWT_HIP(hipMemcpyAsync(s.h_results, s.d_results, nk * 8, hipMemcpyDeviceToHost, s.stream));
  u64 n_ids = 0;
  if (colors_) {
    n_ids = s.colors->pseudoalign_device(s.d_results, s.searcher->kmer_ends_device(), nr, cfg_.color,
                                         s.d_offsets, s.d_ids, s.ids_capacity, nullptr);
    SBWT_HIP(hipMemcpyAsync(s.h_offsets, s.d_offsets, (nr + 1) * 8, hipMemcpyDeviceToHost, s.stream));
    if (n_ids)
      SBWT_HIP(hipMemcpyAsync(s.h_ids, s.d_ids, n_ids * 4, hipMemcpyDeviceToHost, s.stream));
  }
  SBWT_HIP(hipStreamSynchronize(s.stream));
  if (nk && cfg_.kmer_results) std::memcpy(r.kmer_results.data() + kmer_out, s.h_results, nk * 8);
  if (colors_) {
    const u64 base = r.color_ids.size();
    r.color_ids.resize(base + n_ids);
    if (n_ids) std::memcpy(r.color_ids.data() + base, s.h_ids, n_ids * 4);
    for (u64 i = 0; i < nr; ++i) r.color_offsets[read0 + i + 1] = base + s.h_offsets[i + 1];
  }
  return nk;
}

// A read longer than the slot: pieces of at most max_chars_per_slot chars that
// overlap by k - 1, so every k-mer lies in exactly one piece, in order (the
// reference carries the last k - 1 chars of a batch into the next,
// ContinuousSequenceFileParser.cpp:82-112). The k-mer results are the pieces'
// in turn. The color counts and the read's found / not-found / invalid counts
// add up over the pieces and are thresholded once, as a read alone in its
// batch is (ContinuousColorResultsPrinter.hpp:235-263).
u64 Server::run_long_read(Slot &s, Request &r, u64 i, u64 kmer_out) {
  const u64 k = index_->k, C = cfg_.max_chars_per_slot;
  const u64 a0 = i ? r.in_ends[i - 1] - r.ends_base : 0, end = r.in_ends[i] - r.ends_base;
  const u64 nc_colors = colors_ ? colors_->num_colors : 0;
  std::vector<u64> counts(nc_colors, 0);
  u64 F = 0, NF = 0, INV = 0, out = kmer_out;
  s.searcher->set_algorithm(kAlgoAuto);
  s.searcher->set_lane_share(1);
  for (u64 a = a0;;) {
    const u64 b = std::min(end, a + C), nc = b - a;
    std::memcpy(s.h_chars, r.in_chars + a, nc);
    s.h_ends[0] = nc;
    SBWT_HIP(hipMemcpyAsync(s.d_chars, s.h_chars, nc, hipMemcpyHostToDevice, s.stream));
    SBWT_HIP(hipMemcpyAsync(s.d_ends, s.h_ends, 8, hipMemcpyHostToDevice, s.stream));
    u64 nk = 0;
    s.searcher->search_reads_device(s.d_chars, nc, s.d_ends, 1, s.d_results, &nk);
    if (nk && cfg_.kmer_results)
      SBWT_HIP(hipMemcpyAsync(s.h_results, s.d_results, nk * 8, hipMemcpyDeviceToHost, s.stream));
    if (colors_ && nk) {
      s.colors->counts_device(s.d_results, s.searcher->kmer_ends_device(), 1, s.d_piece_counts);
      (void)s.colors->pseudoalign_device(s.d_results, s.searcher->kmer_ends_device(), 1, cfg_.color,
                                         s.d_offsets, s.d_ids, s.ids_capacity, s.d_piece_stats);
      SBWT_HIP(hipMemcpyAsync(s.h_piece, s.d_piece_counts, nc_colors * 4, hipMemcpyDeviceToHost,
                              s.stream));
      SBWT_HIP(hipMemcpyAsync(s.h_piece + nc_colors, s.d_piece_stats, 12, hipMemcpyDeviceToHost,
                              s.stream));
    }
    SBWT_HIP(hipStreamSynchronize(s.stream));
    if (nk && cfg_.kmer_results) std::memcpy(r.kmer_results.data() + out, s.h_results, nk * 8);
    out += nk;
    if (colors_ && nk) {
      for (u64 c = 0; c < nc_colors; ++c) counts[c] += s.h_piece[c];
      F += s.h_piece[nc_colors];
      NF += s.h_piece[nc_colors + 1];
      INV += s.h_piece[nc_colors + 2];
    }
    if (b == end) break;
    a = b - (k - 1);
  }
  if (colors_) {
    const ColorParams &p = cfg_.color;
    if (p.first_read_quirk) {  // the first read of its batch (u64 wrap, as on the device)
      F -= 1;
      NF -= 1;
      INV -= 1;
    }
    const u64 seq = F + (p.include_not_found ? NF : 0) + (p.include_invalid ? INV : 0);
    const u64 min_hits = static_cast<u64>(std::ceil(static_cast<double>(seq) * p.threshold));
    if (min_hits > 0)
      for (u64 c = 0; c < nc_colors; ++c)
        if (counts[c] >= min_hits) r.color_ids.push_back(static_cast<u32>(c));
    r.color_offsets[i + 1] = r.color_ids.size();
  }
  return out - kmer_out;
}

// The batch is padded to its bucket with empty reads (they have no k-mers and
// no color ids); chars past nc in the staging buffer belong to no read.
u64 Server::run_batch_graph(Slot &s, Request &r, u64 nc, u64 nr, u64 read0, u64 kmer_out, u64 cb,
                            u64 rb) {
  for (u64 i = nr; i < rb; ++i) s.h_ends[i] = nc;
  u64 nk = 0;
  const u64 k = index_->k;
  for (u64 i = 0, st = 0; i < nr; ++i) {
    const u64 len = s.h_ends[i] - st;
    nk += len >= k ? len - k + 1 : 0;
    st = s.h_ends[i];
  }
  SBWT_HIP(hipGraphLaunch(graph_for(s, cb, rb), s.stream));
  SBWT_HIP(hipStreamSynchronize(s.stream));
  if (nk && cfg_.kmer_results) std::memcpy(r.kmer_results.data() + kmer_out, s.h_results, nk * 8);
  if (colors_) {
    const u64 n_ids = s.h_offsets[nr];
    if (n_ids > s.ids_capacity)
      throw std::runtime_error("color id output needs " + std::to_string(n_ids) +
                               " entries, capacity " + std::to_string(s.ids_capacity));
    const u64 base = r.color_ids.size();
    r.color_ids.resize(base + n_ids);
    if (n_ids) {
      SBWT_HIP(hipMemcpyAsync(s.h_ids, s.d_ids, n_ids * 4, hipMemcpyDeviceToHost, s.stream));
      SBWT_HIP(hipStreamSynchronize(s.stream));
      std::memcpy(r.color_ids.data() + base, s.h_ids, n_ids * 4);
    }
    for (u64 i = 0; i < nr; ++i) r.color_offsets[read0 + i + 1] = base + s.h_offsets[i + 1];
  }
  return nk;
}

namespace {

// Word-wise 64-bit hash of a request's outputs (multiply-xor per word, then
// a final avalanche); equal outputs give equal digests.
struct Digest {
  u64 h = 0x9E3779B97F4A7C15ULL;
  void add(u64 w) {
    h ^= w;
    h *= 0xFF51AFD7ED558CCDULL;
    h ^= h >> 29;
  }
  u64 done() const {
    u64 x = h ^ (h >> 33);
    x *= 0xC4CEB9FE1A85EC53ULL;
    return x ^ (x >> 33);
  }
};

}  // namespace

void Server::replay(const char *chars, const u64 *read_ends, const u64 *first_read, u64 n_requests,
                    const double *arrival_s, double *latency_s, double *lag_s, u64 *digest,
                    double *wall_s) {
  using clock = std::chrono::steady_clock;
  std::vector<u64> ids(n_requests, 0);
  u64 submitted = 0, limit = n_requests;  // limit: requests that will be submitted
  std::mutex m;
  std::condition_variable cv;
  std::string first_error;
  const clock::time_point t0 = clock::now();
  const auto at = [&](double sec) {
    return t0 + std::chrono::duration_cast<clock::duration>(std::chrono::duration<double>(sec));
  };
  std::thread waiter([&] {
    for (u64 i = 0; i < n_requests; ++i) {
      u64 id = 0;
      {
        std::unique_lock<std::mutex> g(m);
        cv.wait(g, [&] { return submitted > i || limit <= i; });
        if (submitted <= i) break;
        id = ids[i];
      }
      auto r = wait(id);
      const clock::time_point end =
          t0_ + std::chrono::duration_cast<clock::duration>(std::chrono::duration<double>(r->end_s));
      latency_s[i] = std::chrono::duration<double>(end - at(arrival_s[i])).count();
      if (!r->error.empty() && first_error.empty()) first_error = r->error;
      if (digest) {
        Digest d;
        for (u64 v : r->kmer_results) d.add(v);
        for (u64 v : r->color_offsets) d.add(v);
        for (u32 v : r->color_ids) d.add(v);
        digest[i] = d.done();
      }
      release(id);
    }
  });
  std::string submit_error;
  try {
    for (u64 i = 0; i < n_requests; ++i) {
      const clock::time_point target = at(arrival_s[i]);
      // sleep to just before the arrival, then spin to it
      std::this_thread::sleep_until(target - std::chrono::microseconds(200));
      while (clock::now() < target) {
      }
      const u64 a = first_read[i], b = first_read[i + 1];
      const u64 base = a ? read_ends[a - 1] : 0;
      const u64 nc = (b > a ? read_ends[b - 1] : base) - base;
      const u64 id = submit_borrowed(chars + base, nc, read_ends + a, b - a, base);
      lag_s[i] = std::chrono::duration<double>(clock::now() - target).count();
      {
        std::lock_guard<std::mutex> g(m);
        ids[i] = id;
        submitted = i + 1;
      }
      cv.notify_one();
    }
  } catch (const std::exception &e) {
    submit_error = e.what();
  }
  {
    std::lock_guard<std::mutex> g(m);
    limit = submitted;
  }
  cv.notify_one();
  waiter.join();
  if (!submit_error.empty()) throw std::runtime_error(submit_error);
  *wall_s = std::chrono::duration<double>(clock::now() - t0).count();
  if (!first_error.empty()) throw std::runtime_error(first_error);
}

// Captures, once per bucket, search algorithm and work-item size, the staging copies, the search, the color stage
// and the copies back on the slot's stream (timing and census are off in the
// slots, so only kernels and copies are recorded).
hipGraphExec_t Server::graph_for(Slot &s, u64 cb, u64 rb) {
  const auto key = std::make_tuple(cb, rb, s.searcher->algorithm(), s.searcher->stream_chunk(cb));
  auto it = s.graphs.find(key);
  if (it != s.graphs.end()) return it->second;
  hipGraph_t graph = nullptr;
  SBWT_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  try {
    SBWT_HIP(hipMemcpyAsync(s.d_chars, s.h_chars, cb, hipMemcpyHostToDevice, s.stream));
    SBWT_HIP(hipMemcpyAsync(s.d_ends, s.h_ends, rb * 8, hipMemcpyHostToDevice, s.stream));
    s.searcher->search_reads_device(s.d_chars, cb, s.d_ends, rb, s.d_results, nullptr);
    if (cfg_.kmer_results)
      SBWT_HIP(hipMemcpyAsync(s.h_results, s.d_results, cb * 8, hipMemcpyDeviceToHost, s.stream));
    if (colors_) {
      s.colors->pseudoalign_device_async(s.d_results, s.searcher->kmer_ends_device(), rb, cfg_.color,
                                         s.d_offsets, s.d_ids, s.ids_capacity, nullptr);
      SBWT_HIP(hipMemcpyAsync(s.h_offsets, s.d_offsets, (rb + 1) * 8, hipMemcpyDeviceToHost,
                              s.stream));
    }
  } catch (...) {
    (void)hipStreamEndCapture(s.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  SBWT_HIP(hipStreamEndCapture(s.stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  SBWT_HIP(e);
  s.graphs[key] = exec;
  return exec;
}

}  // namespace sbwt_amd
