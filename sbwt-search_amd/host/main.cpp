// sbwt_search: the drop-in command line of the MI355X engine.
//   sbwt_search index  -q <fasta/q(.gz)|.list> -i <.sbwt|.tdbg> -o <prefix|.list>
//                      [-k <.tcolors>] [-p ascii|binary|bool] [--no-headers] [-s streams]
//                      [--batch-chars N] [--host-parse]
//                      [--search-algorithm auto|streaming|per-kmer|adaptive] [--kmer-table]
//   sbwt_search colors -q <index results|.list> -k <.tcolors> -o <prefix|.list>
//                      [-p ascii|binary|csv] [-t threshold] [--include-not-found]
//                      [--include-invalid] [--no-headers] [-s streams] [--batch-bytes N]
// Option names and defaults follow src/ArgumentParser/IndexSearchArgumentParser.cpp:29-178
// and ColorSearchArgumentParser.cpp:27-180; the memory options (-u -m -c -g -r)
// are accepted and bound the batch sizes.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sbwt_amd.h"
#include "io.hpp"

using sbwt_host::u32;
using sbwt_host::u64;
using u8 = uint8_t;

namespace {

struct Args {
  std::string mode;
  std::map<std::string, std::string> kv;
  std::map<std::string, bool> flags;
  std::string get(const std::string &k, const std::string &def = "") const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
  }
  bool flag(const std::string &k) const { return flags.count(k) > 0; }
};

const std::map<std::string, std::string> kShort = {
    {"q", "query-file"},          {"i", "index-file"},      {"o", "output-prefix"},
    {"u", "unavailable-main-memory"}, {"m", "max-main-memory"}, {"c", "cpu-memory-percentage"},
    {"r", "base-pairs-per-seq"},  {"g", "gpu-memory-percentage"}, {"s", "streams"},
    {"p", "print-mode"},          {"k", "colors-file"},     {"t", "threshold"},
    {"h", "help"}};
const char *kFlags[] = {"no-headers", "include-not-found", "include-invalid", "host-parse",
                        "kmer-table", "help"};

Args parse(int argc, char **argv) {
  Args a;
  if (argc < 2) throw std::runtime_error("usage: sbwt_search index|colors [options]");
  a.mode = argv[1];
  for (int i = 2; i < argc; ++i) {
    std::string s = argv[i], key;
    if (s.rfind("--", 0) == 0) key = s.substr(2);
    else if (s.size() == 2 && s[0] == '-' && kShort.count(s.substr(1))) key = kShort.at(s.substr(1));
    else continue;  // unrecognised options are allowed, as in the reference
    std::string val;
    const size_t eq = key.find('=');
    if (eq != std::string::npos) {
      val = key.substr(eq + 1);
      key = key.substr(0, eq);
      a.kv[key] = val;
      continue;
    }
    bool is_flag = false;
    for (const char *f : kFlags) is_flag |= key == f;
    if (is_flag) {
      a.flags[key] = true;
    } else if (i + 1 < argc) {
      a.kv[key] = argv[++i];
    }
  }
  return a;
}

void check(int rc) {
  if (rc != 0) throw std::runtime_error(sbwt_last_error());
}

double fraction(const Args &a, const std::string &key, const std::string &def, const char *msg) {
  const double v = std::stod(a.get(key, def));
  if (v < 0 || v > 1) {
    std::fprintf(stderr, "%s\n", msg);
    std::exit(1);
  }
  return v;
}

// chars per index-search batch from free device memory: input, packed
// sequence, invalid mask, results and offsets (about 13 bytes per char), the
// formatted output (up to 21 bytes per k-mer in ascii) and the device parsers
// (one FASTA parser, about 3 bytes per char, and three FASTQ parsers, about 4
// each), at most 2^24.
// The cap is measured, not a memory limit: file to file, the command is bound
// by writing its output, and with 2^27-char batches (output buffers near
// 1 GB) the writes ran at half the speed they reach with 2^24-char batches,
// while the device stage stays well below the write time at either size.
u64 chars_per_batch(double gpu_fraction, u64 streams) {
  u64 free_b = 0, total = 0;
  check(sbwt_mem_info(&free_b, &total));
  const u64 budget = static_cast<u64>(static_cast<double>(free_b) * gpu_fraction) / std::max<u64>(streams, 1);
  const u64 n = std::min<u64>(budget / 49, 1ULL << 24);
  if (n == 0) throw std::runtime_error("Not enough memory");
  return n;
}

// Busy time of each stage of the index command (seconds summed over
// streams), printed to stderr when SBWT_SEARCH_TIMING is set.
struct StageTimes {
  std::atomic<u64> parse_ns{0}, search_ns{0}, format_ns{0};
};
StageTimes g_times;

struct Clock {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  u64 lap() {
    const auto t = std::chrono::steady_clock::now();
    const u64 ns = static_cast<u64>(std::chrono::duration_cast<std::chrono::nanoseconds>(t - t0).count());
    t0 = t;
    return ns;
  }
};

// Blocking queue of bounded size between the stages of one stream; abort()
// wakes every waiter (a failing stage stops the others).
template <typename T>
class Channel {
 public:
  explicit Channel(size_t cap) : cap_(cap) {}
  bool push(T v) {
    std::unique_lock<std::mutex> g(mu_);
    cv_push_.wait(g, [&] { return q_.size() < cap_ || aborted_; });
    if (aborted_) return false;
    q_.push_back(std::move(v));
    cv_pop_.notify_one();
    return true;
  }
  bool pop(T &v) {
    std::unique_lock<std::mutex> g(mu_);
    cv_pop_.wait(g, [&] { return !q_.empty() || closed_ || aborted_; });
    if (aborted_ || q_.empty()) return false;
    v = std::move(q_.front());
    q_.pop_front();
    cv_push_.notify_one();
    return true;
  }
  void close() {
    std::lock_guard<std::mutex> g(mu_);
    closed_ = true;
    cv_pop_.notify_all();
  }
  void abort() {
    std::lock_guard<std::mutex> g(mu_);
    aborted_ = true;
    cv_pop_.notify_all();
    cv_push_.notify_all();
  }
  // Items left in the channel (after the stages have stopped).
  std::deque<T> drain() {
    std::lock_guard<std::mutex> g(mu_);
    return std::move(q_);
  }

 private:
  size_t cap_;
  std::mutex mu_;
  std::condition_variable cv_push_, cv_pop_;
  std::deque<T> q_;
  bool closed_ = false, aborted_ = false;
};

// Page-locked host buffer (copies to and from the device at link speed).
struct Pinned {
  char *p = nullptr;
  u64 cap = 0;
  void ensure(u64 n) {
    if (n <= cap) return;
    release();
    void *q = nullptr;
    check(sbwt_host_alloc(n, &q));
    p = static_cast<char *>(q);
    cap = n;
  }
  void release() {
    sbwt_host_free(p);
    p = nullptr;
    cap = 0;
  }
};

// One device batch: whole reads, or pieces of a long read (closes[e] == 0:
// the next entry continues the same read, and so the same output line).
// A raw batch instead holds FASTA bytes from a record start to a record end
// (n_chars bytes), parsed on the device; a FASTQ batch was parsed on the
// device by the reader already and is held by its parser (fq, no chars).
struct InBatch {
  u64 file = 0;
  bool first = false, last = false;  // first / last batch of its file
  bool raw = false;
  sbwt_fastq_parser *fq = nullptr;
  Pinned chars;
  u64 n_chars = 0;
  std::vector<u64> ends;
  std::vector<u8> closes;
};

struct OutChunk {
  u64 file = 0;
  bool first = false, last = false;
  Pinned bytes;
  u64 n = 0;
};

// One stream of the index command, three stages on three threads: the reader
// cuts a file into batches; the device stage searches a batch and formats
// its results into the output bytes on the device; the writer appends them
// to the file. Parsing, device work and writing of different batches overlap.
// A FASTA file goes to the device as raw chunks cut at record starts, parsed
// there; from the first chunk holding a line that starts with '+' or a record
// longer than a batch, the rest of the file is parsed by the host record
// reader into batches of reads. A FASTQ file (its first chunk starts with '@'
// and holds a '+' line start) is parsed on the device by the reader thread as
// four-line records, chunk by chunk, up to the first record that the host
// reader would read differently (or one longer than a batch); the host reader
// takes the rest.
void index_stream(sbwt_index *index, u32 k, bool move_to_key, const std::vector<std::string> &in,
                  const std::vector<std::string> &out, const std::vector<u64> &files, int mode,
                  bool headers, u64 max_chars, bool device_parse, int algorithm) {
  const u64 max_reads = std::max<u64>(max_chars / 16, 1024);
  constexpr size_t kBuffers = 3;  // per direction: one filling, one in the queue, one in use
  Channel<InBatch> to_device(1);
  Channel<OutChunk> to_writer(1);
  Channel<Pinned> free_in(kBuffers), free_out(kBuffers);
  Channel<sbwt_fastq_parser *> free_fq(kBuffers);  // FASTQ parsers back from the device stage
  size_t fq_made = 0;
  for (size_t i = 0; i < kBuffers; ++i) {
    Pinned in_buf;
    in_buf.ensure(max_chars + 64);
    free_in.push(in_buf);
    free_out.push(Pinned());
  }
  std::string error;
  std::mutex error_mu;
  auto fail = [&](const std::exception &e) {
    {
      std::lock_guard<std::mutex> g(error_mu);
      if (error.empty()) error = e.what();
    }
    to_device.abort();
    to_writer.abort();
    free_in.abort();
    free_out.abort();
    free_fq.abort();
  };

  std::thread device([&] {
    sbwt_searcher *s = nullptr;
    sbwt_formatter *fmt = nullptr;
    sbwt_fasta_parser *fp = nullptr;
    try {
      check(sbwt_searcher_create(index, max_chars, max_reads, move_to_key, nullptr, &s));
      check(sbwt_searcher_set_algorithm(s, algorithm));
      check(sbwt_formatter_create(max_chars, max_reads, nullptr, &fmt));
      bool line_open = false;
      InBatch b;
      while (to_device.pop(b)) {
        Clock clk;
        OutChunk o;
        o.file = b.file;
        o.first = b.first;
        o.last = b.last;
        if (!free_out.pop(o.bytes)) break;
        if (b.first) line_open = false;
        if (b.raw && b.n_chars) {
          if (!fp) check(sbwt_fasta_parser_create(max_chars, max_reads, nullptr, &fp));
          u64 nr = 0;
          check(sbwt_searcher_search_fasta_format(s, fp, fmt, mode, b.chars.p, b.n_chars,
                                                  o.bytes.p, o.bytes.cap, &o.n, &nr));
          if (o.n > o.bytes.cap) {
            o.bytes.ensure(o.n + o.n / 4);
            check(sbwt_formatter_copy_last(fmt, o.bytes.p, o.bytes.cap));
          }
          line_open = false;
        } else if (b.fq) {
          check(sbwt_searcher_search_fastq_format(s, b.fq, fmt, mode, o.bytes.p, o.bytes.cap, &o.n));
          if (o.n > o.bytes.cap) {
            o.bytes.ensure(o.n + o.n / 4);
            check(sbwt_formatter_copy_last(fmt, o.bytes.p, o.bytes.cap));
          }
          line_open = false;
          sbwt_fastq_parser *q = b.fq;
          b.fq = nullptr;
          if (!free_fq.push(q)) {
            sbwt_fastq_parser_destroy(q);
            break;
          }
        } else if (!b.raw && !b.ends.empty()) {
          check(sbwt_searcher_search_reads_format(s, fmt, mode, b.chars.p, b.n_chars,
                                                  b.ends.data(), b.closes.data(), b.ends.size(),
                                                  line_open, o.bytes.p, o.bytes.cap, &o.n));
          if (o.n > o.bytes.cap) {
            o.bytes.ensure(o.n + o.n / 4);
            check(sbwt_formatter_copy_last(fmt, o.bytes.p, o.bytes.cap));
          }
          line_open = b.closes.back() == 0;
        }
        if (b.chars.p && !free_in.push(b.chars)) break;
        g_times.search_ns += clk.lap();
        if (!to_writer.push(o)) break;
      }
      to_writer.close();
    } catch (const std::exception &e) {
      fail(e);
    }
    if (fp) sbwt_fasta_parser_destroy(fp);
    if (fmt) sbwt_formatter_destroy(fmt);
    if (s) sbwt_searcher_destroy(s);
  });

  std::thread writer([&] {
    try {
      std::unique_ptr<sbwt_host::IndexFileOut> f;
      OutChunk o;
      while (to_writer.pop(o)) {
        Clock clk;
        if (o.first) f = std::make_unique<sbwt_host::IndexFileOut>(out[o.file], mode, headers);
        if (o.n) f->write(o.bytes.p, o.n);
        if (o.last) f.reset();
        g_times.format_ns += clk.lap();
        if (!free_out.push(o.bytes)) break;
      }
    } catch (const std::exception &e) {
      fail(e);
    }
  });

  // reader (this thread)
  try {
    std::vector<char> read;
    for (size_t fi = 0; fi < files.size(); ++fi) {
      const u64 f = files[fi];
      std::unique_ptr<sbwt_host::SeqReader> reader;
      std::unique_ptr<sbwt_host::RawSeqFile> raw;
      try {
        if (device_parse) raw = std::make_unique<sbwt_host::RawSeqFile>(in[f]);
        else reader = std::make_unique<sbwt_host::SeqReader>(in[f]);
      } catch (const sbwt_host::InputMissing &e) {
        std::fprintf(stderr, "%s\n", e.what());  // the file's output holds headers only
      }
      InBatch b;
      bool ok = free_in.pop(b.chars);
      b.file = f;
      b.first = true;
      Clock clk;
      if (ok && raw) {
        // Raw chunks: b.chars[0, n_chars) holds bytes from a record start on.
        // Bytes before the first '>' or '@' are skipped, as the record reader
        // does.
        bool started = false;
        int kind = -1;  // 0 FASTA, 1 FASTQ: decided on the first chunk
        while (ok) {
          const u64 got = raw->read(b.chars.p + b.n_chars, max_chars - b.n_chars);
          if (!started) {
            const char *p = b.chars.p;
            const char *h1 = static_cast<const char *>(std::memchr(p, '>', got));
            const char *h2 = static_cast<const char *>(std::memchr(p, '@', got));
            const char *h = h1 && h2 ? std::min(h1, h2) : (h1 ? h1 : h2);
            if (!h) {
              if (raw->at_end()) break;  // no record in the file
              continue;                  // all of it skipped; read on
            }
            std::memmove(b.chars.p, h, static_cast<size_t>(p + got - h));
            b.n_chars = static_cast<u64>(p + got - h);
            started = true;
            if (b.n_chars < max_chars && !raw->at_end()) continue;
          } else {
            b.n_chars += got;
          }
          const bool last = raw->at_end();
          if (kind < 0)
            kind = b.chars.p[0] == '@' && sbwt_fasta_has_quality_line(b.chars.p, b.n_chars, 1) ? 1 : 0;
          if (kind == 1) {
            sbwt_fastq_parser *q = nullptr;
            if (fq_made < kBuffers) {
              check(sbwt_fastq_parser_create(max_chars, max_reads, nullptr, &q));
              ++fq_made;
            } else if (!free_fq.pop(q)) {
              ok = false;
              break;
            }
            u64 used = 0, nc = 0, nr = 0;
            const int rc = sbwt_fastq_parser_parse(q, b.chars.p, b.n_chars, last ? 1 : 0, nullptr,
                                                   0, nullptr, 0, &used, &nc, &nr);
            if (rc != 0 || used == 0) {
              free_fq.push(q);
              check(rc);
              // the rest of the file goes to the host reader
              reader = raw->to_reader(b.chars.p, b.n_chars);
              b.n_chars = 0;
              break;
            }
            InBatch pb;
            pb.file = f;
            pb.first = b.first;
            pb.fq = q;
            b.first = false;
            const u64 tail = b.n_chars - used;
            if (tail) std::memmove(b.chars.p, b.chars.p + used, tail);
            b.n_chars = tail;
            g_times.parse_ns += clk.lap();
            if (!to_device.push(std::move(pb))) {
              sbwt_fastq_parser_destroy(q);
              ok = false;
              break;
            }
            if (last && tail == 0) break;
            continue;
          }
          const u64 cut = sbwt_fasta_has_quality_line(b.chars.p, b.n_chars, 1)
                              ? 0
                              : sbwt_fasta_chunk_cut(b.chars.p, b.n_chars, last, max_reads);
          if (cut == 0) {  // the rest of the file goes to the host reader
            reader = raw->to_reader(b.chars.p, b.n_chars);
            b.n_chars = 0;
            break;
          }
          Pinned next;
          if (!free_in.pop(next)) {
            ok = false;
            break;
          }
          const u64 tail = b.n_chars - cut;
          if (tail) std::memcpy(next.p, b.chars.p + cut, tail);
          b.n_chars = cut;
          b.raw = true;
          g_times.parse_ns += clk.lap();
          if (!to_device.push(std::move(b))) {
            next.release();
            ok = false;
            break;
          }
          b = InBatch();
          b.file = f;
          b.chars = next;
          b.n_chars = tail;
          if (last && tail == 0) break;
        }
      }
      // hands the batch to the device stage; a batch that is not the file's
      // last one is followed by a fresh buffer from the pool
      auto flush_batch = [&](bool last) {
        b.last = last;
        g_times.parse_ns += clk.lap();
        if (!to_device.push(std::move(b))) return false;
        b = InBatch();
        b.file = f;
        return last || free_in.pop(b.chars);
      };
      while (ok) {
        read.clear();
        if (!reader || !reader->next(read)) break;
        // a read longer than the room left is cut into pieces overlapping by
        // k-1 chars (the carry-over of ContinuousSequenceFileParser.cpp:82-112)
        u64 st = 0;
        while (ok) {
          const u64 room = max_chars - b.n_chars;
          if (room < k && !b.ends.empty()) {
            ok = flush_batch(false);
            continue;
          }
          const u64 take = std::min<u64>(read.size() - st, room);
          if (take) std::memcpy(b.chars.p + b.n_chars, read.data() + st, take);
          b.n_chars += take;
          b.ends.push_back(b.n_chars);
          const bool last = st + take == read.size();
          b.closes.push_back(last ? 1 : 0);
          if (last) break;
          st += take - (k - 1);
          ok = flush_batch(false);
        }
        if (ok && (b.n_chars >= max_chars || b.ends.size() >= max_reads)) ok = flush_batch(false);
      }
      if (ok) ok = flush_batch(true);  // the file's last batch (possibly empty)
      else if (b.chars.p) b.chars.release();  // stopped by a failure elsewhere
      if (!ok) break;
    }
    to_device.close();
  } catch (const std::exception &e) {
    fail(e);
  }
  device.join();
  writer.join();
  // every buffer is back in its pool (or stranded in a queue after a failure)
  for (auto &x : free_in.drain()) x.release();
  for (auto &x : free_out.drain()) x.release();
  for (auto &x : to_device.drain()) {
    x.chars.release();
    if (x.fq) sbwt_fastq_parser_destroy(x.fq);
  }
  for (auto &x : to_writer.drain()) x.bytes.release();
  for (auto *q : free_fq.drain()) sbwt_fastq_parser_destroy(q);
  if (!error.empty()) throw std::runtime_error(error);
}

int run_index(const Args &a) {
  const std::string q = a.get("query-file"), idx = a.get("index-file"), o = a.get("output-prefix");
  if (q.empty() || idx.empty() || o.empty())
    throw std::runtime_error("query-file, index-file and output-prefix are required");
  const std::string colors = a.get("colors-file");
  const std::string mode_name = a.get("print-mode", "ascii");
  if (mode_name != "ascii" && mode_name != "binary" && mode_name != "bool")
    throw std::runtime_error("Invalid value passed by user for argument print_mode");
  const int mode = mode_name == "ascii" ? SBWT_FORMAT_ASCII
                   : mode_name == "binary" ? SBWT_FORMAT_BINARY : SBWT_FORMAT_BOOL;
  fraction(a, "cpu-memory-percentage", "0.8", "Invalid value for cpu-memory-percentage. Must be between 0 and 1.");
  const double gpu = fraction(a, "gpu-memory-percentage", "0.95",
                              "Invalid value for gpu-memory-percentage. Must be between 0 and 1.");
  const auto in = sbwt_host::expand_list(q), out = sbwt_host::expand_list(o);
  if (in.size() != out.size()) throw std::runtime_error("Input and output file sizes differ");
  const u64 streams = std::max<u64>(1, std::min<u64>(in.size(), std::stoull(a.get("streams", "4"))));
  sbwt_index *index = nullptr;
  check(sbwt_index_load(idx.c_str(), colors.empty() ? nullptr : colors.c_str(), 0, &index));
  sbwt_index_info info;
  check(sbwt_index_get_info(index, &info));
  u64 max_chars = chars_per_batch(gpu, streams);
  // --batch-chars N (an extension) caps the chars per device batch
  if (!a.get("batch-chars").empty())
    max_chars = std::max<u64>(std::min<u64>(max_chars, std::stoull(a.get("batch-chars"))), info.kmer_size);
  // --host-parse (an extension) keeps every sequence file on the host reader
  const bool device_parse = !a.flag("host-parse");
  // --search-algorithm (an extension): the search kernel of every batch; the
  // default picks by batch size. Results are identical for all four.
  const std::string algo_name = a.get("search-algorithm", "auto");
  const std::map<std::string, int> algos = {{"auto", SBWT_ALGO_AUTO},
                                            {"streaming", SBWT_ALGO_STREAMING},
                                            {"per-kmer", SBWT_ALGO_PER_KMER},
                                            {"adaptive", SBWT_ALGO_STREAMING_ADAPTIVE}};
  if (!algos.count(algo_name)) throw std::runtime_error("Invalid value passed by user for argument search-algorithm");
  const int algorithm = algos.at(algo_name);
  // --kmer-table (an extension): build the index's k-mer table before searching
  if (a.flag("kmer-table")) {
    int built = 0;
    check(sbwt_index_build_kmer_table(index, &built));
  }
  const auto parts = sbwt_host::partition_by_size(in, streams);
  std::vector<std::thread> th;
  std::vector<std::string> errors(streams);
  for (u64 p = 0; p < streams; ++p)
    th.emplace_back([&, p] {
      try {
        index_stream(index, info.kmer_size, !colors.empty(), in, out, parts[p], mode,
                     !a.flag("no-headers"), max_chars, device_parse, algorithm);
      } catch (const std::exception &e) {
        errors[p] = e.what();
      }
    });
  for (auto &t : th) t.join();
  sbwt_index_destroy(index);
  for (auto &e : errors)
    if (!e.empty()) throw std::runtime_error(e);
  if (std::getenv("SBWT_SEARCH_TIMING"))
    std::fprintf(stderr, "stage busy seconds (summed over %llu streams; stages overlap): parse %.3f search+format %.3f write %.3f\n",
                 static_cast<unsigned long long>(streams), g_times.parse_ns * 1e-9,
                 g_times.search_ns * 1e-9, g_times.format_ns * 1e-9);
  return 0;
}

// --------------------------------------------------------------- colors
// One colors batch: body bytes of one index-result file that end at a read
// end (or at the end of the file).
struct ColorInBatch {
  u64 file = 0;
  bool first = false, last = false;  // first / last batch of its file
  int format = 0;
  Pinned body;
  u64 n = 0;
};

// Whether bytes of an index-file body may hold a read end: a '\n' in ascii
// (not every '\n' is one), a u64 max - 2 word in binary.
bool holds_read_end(int format, const char *p, u64 n) {
  if (format == SBWT_INDEX_FILE_ASCII) return std::memchr(p, '\n', n) != nullptr;
  for (u64 w = 0; w + 8 <= n; w += 8) {
    u64 v;
    std::memcpy(&v, p + w, 8);
    if (v == ~0ULL - 2) return true;
  }
  return false;
}

// One stream of the colors command, three stages on three threads: the reader
// cuts index-file bodies into batches at read ends; the device stage parses a
// batch, runs the color search and prints the color ids into the output
// bytes, all on the device; the writer appends the bytes to the file.
void colors_stream(sbwt_color_index *ci, u64 num_colors, const std::vector<std::string> &in,
                   const std::vector<std::string> &out, const std::vector<u64> &files, int mode,
                   bool headers, const sbwt_color_params &params, u64 max_bytes, u64 max_reads) {
  constexpr size_t kBuffers = 3;
  Channel<ColorInBatch> to_device(1);
  Channel<OutChunk> to_writer(1);
  Channel<Pinned> free_in(kBuffers), free_out(kBuffers);
  for (size_t i = 0; i < kBuffers; ++i) {
    Pinned in_buf;
    in_buf.ensure(max_bytes + 64);
    free_in.push(in_buf);
    free_out.push(Pinned());
  }
  std::string error;
  std::mutex error_mu;
  auto fail = [&](const std::exception &e) {
    {
      std::lock_guard<std::mutex> g(error_mu);
      if (error.empty()) error = e.what();
    }
    to_device.abort();
    to_writer.abort();
    free_in.abort();
    free_out.abort();
  };

  std::thread device([&] {
    sbwt_color_file_searcher *fs = nullptr;
    try {
      check(sbwt_color_file_searcher_create(ci, max_bytes, max_reads, nullptr, &fs));
      // the reference's printer carries u64 max into the first read of the
      // stream (ContinuousColorResultsPrinter.hpp:62-64,175-180)
      bool first_read = true;
      ColorInBatch b;
      while (to_device.pop(b)) {
        Clock clk;
        OutChunk o;
        o.file = b.file;
        o.first = b.first;
        o.last = b.last;
        if (!free_out.pop(o.bytes)) break;
        sbwt_color_params p = params;
        p.first_read_quirk = first_read ? 1 : 0;
        u64 reads = 0;
        const u64 want = std::max<u64>(b.n * 2, 1 << 16);
        o.bytes.ensure(want);
        check(sbwt_color_file_searcher_batch(fs, b.format, b.body.p, b.n, &p, mode, o.bytes.p,
                                             o.bytes.cap, &o.n, &reads));
        if (o.n > o.bytes.cap) {
          o.bytes.ensure(o.n + o.n / 4);
          check(sbwt_color_file_searcher_copy_last(fs, o.bytes.p, o.bytes.cap));
        }
        if (reads) first_read = false;
        if (!free_in.push(b.body)) break;
        g_times.search_ns += clk.lap();
        if (!to_writer.push(o)) break;
      }
      to_writer.close();
    } catch (const std::exception &e) {
      fail(e);
    }
    if (fs) sbwt_color_file_searcher_destroy(fs);
  });

  std::thread writer([&] {
    try {
      std::unique_ptr<sbwt_host::ColorFileOut> f;
      OutChunk o;
      while (to_writer.pop(o)) {
        Clock clk;
        if (o.first) f = std::make_unique<sbwt_host::ColorFileOut>(out[o.file], mode, headers, num_colors);
        if (o.n) f->write(o.bytes.p, o.n);
        if (o.last) f.reset();
        g_times.format_ns += clk.lap();
        if (!free_out.push(o.bytes)) break;
      }
    } catch (const std::exception &e) {
      fail(e);
    }
  });

  // reader (this thread)
  try {
    bool ok = true;
    for (size_t fi = 0; ok && fi < files.size(); ++fi) {
      const u64 f = files[fi];
      std::unique_ptr<sbwt_host::IndexFileIn> file;
      try {
        file = std::make_unique<sbwt_host::IndexFileIn>(in[f]);
      } catch (const sbwt_host::InputMissing &e) {
        std::fprintf(stderr, "%s\n", e.what());  // the file's output holds headers only
      }
      ColorInBatch b;
      ok = free_in.pop(b.body);
      b.file = f;
      b.first = true;
      b.format = file ? file->format() : SBWT_INDEX_FILE_ASCII;
      bool eof = !file;
      Clock clk;
      while (ok) {
        if (!eof) {
          const u64 got = file->read(b.body.p + b.n, max_bytes - b.n);
          eof = b.n + got < max_bytes;
          b.n += got;
        }
        u64 cut = sbwt_index_file_batch_cut(b.format, b.body.p, b.n, max_reads);
        if (eof) {
          // the rest is the file's last batch unless it holds more read
          // ends than a batch may (then cut after the first max_reads)
          const bool more = cut > 0 && cut < b.n && holds_read_end(b.format, b.body.p + cut, b.n - cut);
          if (!more) {
            b.last = true;
            g_times.parse_ns += clk.lap();
            ok = to_device.push(std::move(b));
            break;
          }
        }
        if (cut == 0)
          throw std::runtime_error("no read end within one batch (" + std::to_string(max_bytes) +
                                   " bytes) of index file " + in[f] + "; raise --batch-bytes");
        // end the batch after a read with a found index, so that reads with
        // none (printed with the next such read's counts under the include
        // flags) stay with it; a batch with no such read keeps its cut
        if (const u64 cc = sbwt_index_file_colored_cut(b.format, b.body.p, cut)) cut = cc;
        ColorInBatch next;
        if (!free_in.pop(next.body)) {
          ok = false;
          break;
        }
        next.file = f;
        next.format = b.format;
        next.n = b.n - cut;
        if (next.n) std::memcpy(next.body.p, b.body.p + cut, next.n);
        b.n = cut;
        g_times.parse_ns += clk.lap();
        Pinned sent = b.body;  // back to the pool through the device stage, or freed here
        ok = to_device.push(std::move(b));
        b = std::move(next);
        if (!ok) sent.release();
      }
      if (!ok && b.body.p) b.body.release();
    }
    to_device.close();
  } catch (const std::exception &e) {
    fail(e);
  }
  device.join();
  writer.join();
  for (auto &x : free_in.drain()) x.release();
  for (auto &x : free_out.drain()) x.release();
  for (auto &x : to_device.drain()) x.body.release();
  for (auto &x : to_writer.drain()) x.bytes.release();
  if (!error.empty()) throw std::runtime_error(error);
}

int run_colors(const Args &a) {
  const std::string q = a.get("query-file"), k = a.get("colors-file"), o = a.get("output-prefix");
  if (q.empty() || k.empty() || o.empty())
    throw std::runtime_error("query-file, colors-file and output-prefix are required");
  const std::string mode_name = a.get("print-mode", "ascii");
  if (mode_name != "ascii" && mode_name != "binary" && mode_name != "csv")
    throw std::runtime_error("Invalid value passed by user for argument print_mode");
  const int mode = mode_name == "ascii" ? SBWT_COLOR_FORMAT_ASCII
                   : mode_name == "binary" ? SBWT_COLOR_FORMAT_BINARY : SBWT_COLOR_FORMAT_CSV;
  const double t = std::stod(a.get("threshold", "1"));
  if (t < 0 || t > 1) {
    std::fprintf(stderr, "Invalid value for threshold, must be between 1 and 0 (both included)\n");
    std::exit(1);
  }
  fraction(a, "cpu-memory-percentage", "0.8", "Invalid value for cpu-memory-percentage. Must be between 0 and 1.");
  const double gpu = fraction(a, "gpu-memory-percentage", "0.95",
                              "Invalid value for gpu-memory-percentage. Must be between 0 and 1.");
  const auto in = sbwt_host::expand_list(q), out = sbwt_host::expand_list(o);
  if (in.size() != out.size()) throw std::runtime_error("Input and output file sizes differ");
  const u64 streams = std::max<u64>(1, std::min<u64>(in.size(), std::stoull(a.get("streams", "4"))));
  sbwt_color_index *ci = nullptr;
  check(sbwt_color_index_load(k.c_str(), &ci));
  sbwt_color_index_info info;
  check(sbwt_color_index_get_info(ci, &info));
  sbwt_color_params params{t, a.flag("include-not-found"), a.flag("include-invalid"), 0};
  // Batch bytes per stream from free device memory: body, results and ids
  // (about 7 bytes per body byte) plus the printer output; reads per batch
  // bounded by the per-read color bitmap.
  u64 free_b = 0, total = 0;
  check(sbwt_mem_info(&free_b, &total));
  const u64 budget = static_cast<u64>(static_cast<double>(free_b) * gpu) / streams;
  u64 max_bytes = std::min<u64>(budget / 16, 64ULL << 20);
  if (!a.get("batch-bytes").empty())  // an extension: caps the bytes per batch
    max_bytes = std::min<u64>(max_bytes, std::stoull(a.get("batch-bytes")));
  if (max_bytes < 16) throw std::runtime_error("Not enough memory");
  const u64 words = (info.num_colors + 63) / 64;
  const u64 max_reads = std::max<u64>(1, std::min<u64>(max_bytes, (256ULL << 20) / (8 * std::max<u64>(words, 1))));
  const auto parts = sbwt_host::partition_by_size(in, streams);
  std::vector<std::thread> th;
  std::vector<std::string> errors(streams);
  for (u64 p = 0; p < streams; ++p)
    th.emplace_back([&, p] {
      try {
        colors_stream(ci, info.num_colors, in, out, parts[p], mode, !a.flag("no-headers"), params,
                      max_bytes, max_reads);
      } catch (const std::exception &e) {
        errors[p] = e.what();
      }
    });
  for (auto &x : th) x.join();
  sbwt_color_index_destroy(ci);
  for (auto &e : errors)
    if (!e.empty()) throw std::runtime_error(e);
  if (std::getenv("SBWT_SEARCH_TIMING"))
    std::fprintf(stderr, "stage busy seconds (summed over %llu streams; stages overlap): read %.3f device %.3f write %.3f\n",
                 static_cast<unsigned long long>(streams), g_times.parse_ns * 1e-9,
                 g_times.search_ns * 1e-9, g_times.format_ns * 1e-9);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    const Args a = parse(argc, argv);
    if (a.mode == "index") return run_index(a);
    if (a.mode == "colors") return run_colors(a);
    std::fprintf(stderr, "usage: sbwt_search index|colors [options]\n");
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
