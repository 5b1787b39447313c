// Host I/O of the sbwt_search command line (see io.hpp).
#include "io.hpp"

#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <sys/stat.h>

namespace sbwt_host {

namespace {
constexpr u64 kNotFound = ~0ULL;
constexpr u64 kInvalid = ~0ULL - 1;
constexpr u64 kNewline = ~0ULL - 2;
constexpr size_t kFlush = 1 << 24;

void write_all(FILE *f, const void *p, size_t n) {
  if (n && fwrite(p, 1, n, f) != n) throw std::runtime_error("write failed");
}
}  // namespace

void write_sized_string(FILE *f, const std::string &s) {
  const u64 n = s.size();
  write_all(f, &n, 8);
  write_all(f, s.data(), n);
}

// ------------------------------------------------------------------ reader
SeqReader::SeqReader(const std::string &path) : buf_(1 << 20) {
  gzFile g = gzopen(path.c_str(), "rb");
  if (!g) throw std::runtime_error("Unable to open file " + path);
  gzbuffer(g, 1 << 20);
  gz_ = g;
}

SeqReader::~SeqReader() {
  if (gz_) gzclose(static_cast<gzFile>(gz_));
}

bool SeqReader::fill() {
  if (pos_ < len_) return true;
  if (eof_) return false;
  const int n = gzread(static_cast<gzFile>(gz_), buf_.data(), static_cast<unsigned>(buf_.size()));
  if (n <= 0) {
    eof_ = true;
    return false;
  }
  len_ = static_cast<size_t>(n);
  pos_ = 0;
  return true;
}

int SeqReader::getc() {
  if (!fill()) return -1;
  return static_cast<unsigned char>(buf_[pos_++]);
}

int SeqReader::peek() {
  if (!fill()) return -1;
  return static_cast<unsigned char>(buf_[pos_]);
}

// Consumes the rest of the line and its '\n'; appends its chars to `out`
// when not null. Returns the number of chars before the '\n'.
u64 SeqReader::take_line(std::vector<char> *out) {
  u64 n = 0;
  while (fill()) {
    const char *p = buf_.data() + pos_;
    const size_t avail = len_ - pos_;
    const void *nl = std::memchr(p, '\n', avail);
    const size_t m = nl ? static_cast<size_t>(static_cast<const char *>(nl) - p) : avail;
    if (out) out->insert(out->end(), p, p + m);
    n += m;
    pos_ += m;
    if (nl) {
      ++pos_;  // the '\n'
      break;
    }
  }
  return n;
}

bool SeqReader::next(std::vector<char> &chars) {
  int c;
  if (!started_) {
    started_ = true;
    while ((c = getc()) != -1 && c != '>' && c != '@') {
    }
    if (c == -1) return false;
  } else if (pending_ == -1) {
    return false;
  }
  pending_ = -1;
  take_line(nullptr);  // header line
  u64 seq_len = 0;
  for (;;) {
    c = peek();
    if (c == -1) break;
    if (c == '>' || c == '@') {
      pending_ = getc();
      break;
    }
    if (c == '+') {  // FASTQ quality: lines until as long as the sequence
      take_line(nullptr);
      u64 q = 0;
      while (q < seq_len && peek() != -1) q += take_line(nullptr);
      while ((c = getc()) != -1 && c != '>' && c != '@') {
      }
      pending_ = c;
      break;
    }
    seq_len += take_line(&chars);
  }
  return true;
}

// ------------------------------------------------------------ index writer
IndexFileOut::IndexFileOut(const std::string &path, int mode, bool headers) {
  static const char *ext[3] = {".txt", ".bin", ".bool"};
  static const char *fmt[3] = {"ascii", "binary", "bool"};
  static const char *ver[3] = {"v1.0", "v1.0", "v2.0"};
  if (mode < 0 || mode > 2) throw std::runtime_error("Invalid value passed by user for argument print_mode");
  f_ = fopen((path + ext[mode]).c_str(), "wb");
  if (!f_) throw std::runtime_error("Unable to open file " + path + ext[mode]);
  if (headers) {
    write_sized_string(f_, fmt[mode]);
    write_sized_string(f_, ver[mode]);
  }
}

IndexFileOut::~IndexFileOut() { fclose(f_); }

void IndexFileOut::write(const char *p, u64 n) { write_all(f_, p, n); }

// ------------------------------------------------------------ color writer
ColorWriter::ColorWriter(const std::string &path, const std::string &mode, bool headers,
                         u64 num_colors)
    : num_colors_(num_colors) {
  static const char *ext[3] = {".txt", ".bin", ".csv"};
  if (mode == "ascii") mode_ = 0;
  else if (mode == "binary") mode_ = 1;
  else if (mode == "csv") mode_ = 2;
  else throw std::runtime_error("Invalid value passed by user for argument print_mode");
  f_ = fopen((path + ext[mode_]).c_str(), "wb");
  if (!f_) throw std::runtime_error("Unable to open file " + path + ext[mode_]);
  if (mode_ == 2) {  // always written (CsvContinuousColorResultsPrinter.cpp:60-75)
    std::string h;
    for (u64 i = 1; i <= num_colors_; ++i) {
      h += std::to_string(i);
      h += i == num_colors_ ? '\n' : ',';
    }
    write_all(f_, h.data(), h.size());
  } else if (headers) {
    write_sized_string(f_, mode_ == 0 ? "ascii" : "binary");
    write_sized_string(f_, "v1.0");
  }
  out_.reserve(kFlush + 64);
}

ColorWriter::~ColorWriter() {
  try {
    flush();
  } catch (...) {
  }
  fclose(f_);
}

void ColorWriter::read(const u32 *ids, u64 n) {
  if (mode_ == 0) {
    for (u64 i = 0; i < n; ++i) {
      if (i) out_.push_back(' ');
      char tmp[16];
      auto r = std::to_chars(tmp, tmp + sizeof tmp, ids[i]);
      out_.insert(out_.end(), tmp, r.ptr);
    }
    out_.push_back('\n');
  } else if (mode_ == 1) {
    for (u64 i = 0; i < n; ++i) {
      const u64 v = ids[i];
      const char *p = reinterpret_cast<const char *>(&v);
      out_.insert(out_.end(), p, p + 8);
    }
    const u64 end = ~0ULL;
    const char *p = reinterpret_cast<const char *>(&end);
    out_.insert(out_.end(), p, p + 8);
  } else {
    const size_t at = out_.size();
    for (u64 c = 0; c < num_colors_; ++c) {
      out_.push_back('0');
      out_.push_back(c + 1 == num_colors_ ? '\n' : ',');
    }
    for (u64 i = 0; i < n; ++i) out_[at + 2 * ids[i]] = '1';
  }
  if (out_.size() >= kFlush) flush();
}

void ColorWriter::flush() {
  write_all(f_, out_.data(), out_.size());
  out_.clear();
}

// ----------------------------------------------------- index file reading
IndexFile read_index_file(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("Unable to open file " + path);
  std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  size_t p = 0;
  auto sized = [&]() {
    u64 n;
    if (p + 8 > data.size()) throw std::runtime_error("Invalid file format in file: " + path);
    std::memcpy(&n, data.data() + p, 8);
    p += 8;
    if (n > 64 || p + n > data.size()) throw std::runtime_error("Invalid file format in file: " + path);
    std::string s(data.data() + p, n);
    p += n;
    return s;
  };
  const std::string fmt = sized();
  const std::string ver = sized();
  if (ver != "v1.0") throw std::runtime_error("The file has an incompatible version number");
  IndexFile out;
  if (fmt == "ascii") {  // AsciiIndexFileParser.cpp:33-70
    u64 n = 0;
    while (p < data.size()) {
      const char c = data[p++];
      if (c == '-') {
        const char d = p < data.size() ? data[p++] : 0;
        if (d == '1') out.results.push_back(kNotFound);
        else if (d == '2') out.results.push_back(kInvalid);
        while (p < data.size() && data[p] >= '0') ++p;
        ++n;
      } else if (c >= '0') {
        u64 v = static_cast<u64>(c - '0');
        while (p < data.size() && data[p] >= '0') v = v * 10 + static_cast<u64>(data[p++] - '0');
        out.results.push_back(v);
        ++n;
      } else if (c == '\n') {
        out.kmer_ends.push_back(out.results.size());
      }
    }
    (void)n;
  } else if (fmt == "binary") {  // BinaryIndexFileParser.cpp:33-63
    while (p + 8 <= data.size()) {
      u64 v;
      std::memcpy(&v, data.data() + p, 8);
      p += 8;
      if (v == kNewline) out.kmer_ends.push_back(out.results.size());
      else out.results.push_back(v);
    }
  } else {
    throw std::runtime_error("Invalid file format in file: " + path);
  }
  return out;
}

std::vector<std::string> expand_list(const std::string &name) {
  const std::string ext = ".list";
  if (name.size() >= ext.size() && name.compare(name.size() - ext.size(), ext.size(), ext) == 0) {
    std::ifstream in(name);
    if (!in) throw std::runtime_error("Unable to open file " + name);
    std::vector<std::string> out;
    std::string line;
    while (std::getline(in, line)) out.push_back(line);
    return out;
  }
  return {name};
}

u64 file_size(const std::string &path) {
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return 0;
  return static_cast<u64>(st.st_size);
}

std::vector<std::vector<u64>> partition_by_size(const std::vector<std::string> &files, u64 parts) {
  std::map<u64, std::vector<u64>> by_size;
  for (u64 i = 0; i < files.size(); ++i) by_size[file_size(files[i])].push_back(i);
  std::vector<std::vector<u64>> out(parts);
  std::vector<u64> load(parts, 0);
  for (auto it = by_size.rbegin(); it != by_size.rend(); ++it) {
    for (u64 f : it->second) {
      const u64 s = std::min_element(load.begin(), load.end()) - load.begin();
      out[s].push_back(f);
      load[s] += it->first;
    }
  }
  return out;
}

}  // namespace sbwt_host
