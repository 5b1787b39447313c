// Host I/O of the sbwt_search command line (see io.hpp).
#include "io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <sys/stat.h>

namespace sbwt_host {

namespace {
// ThrowingOfstream's message (src/Tools/IOUtils.cpp:27-37)
std::string output_open_error(const std::string &path) {
  return "The path " + path +
         " cannot be opened. Check that all the folders in the path is correct and that you "
         "have permission to create files in this path folder";
}

// Large batches go out in 16 MiB writes.
void write_all(FILE *f, const void *p, size_t n) {
  constexpr size_t kPiece = size_t{16} << 20;
  const char *c = static_cast<const char *>(p);
  for (size_t o = 0; o < n; o += kPiece) {
    const size_t m = std::min(kPiece, n - o);
    if (fwrite(c + o, 1, m, f) != m) throw std::runtime_error("write failed");
  }
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
  if (!g) throw InputMissing("The input file " + path + " cannot be opened");
  gzbuffer(g, 1 << 20);
  gz_ = g;
}

SeqReader::SeqReader(void *gz, const char *pending, u64 n)
    : gz_(gz), buf_(std::max<u64>(n, 1 << 20)) {
  if (n) std::memcpy(buf_.data(), pending, n);
  len_ = static_cast<size_t>(n);
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

RawSeqFile::RawSeqFile(const std::string &path) {
  gzFile g = gzopen(path.c_str(), "rb");
  if (!g) throw InputMissing("The input file " + path + " cannot be opened");
  gzbuffer(g, 1 << 20);
  gz_ = g;
}

RawSeqFile::~RawSeqFile() {
  if (gz_) gzclose(static_cast<gzFile>(gz_));
}

u64 RawSeqFile::read(char *dst, u64 n) {
  u64 got = 0;
  while (got < n && !eof_) {
    const unsigned ask = static_cast<unsigned>(std::min<u64>(n - got, 1u << 30));
    const int r = gzread(static_cast<gzFile>(gz_), dst + got, ask);
    if (r < 0) throw std::runtime_error("read error in a sequence file");
    if (r == 0) eof_ = true;
    got += static_cast<u64>(r);
  }
  return got;
}

std::unique_ptr<SeqReader> RawSeqFile::to_reader(const char *pending, u64 n) {
  auto r = std::make_unique<SeqReader>(gz_, pending, n);
  gz_ = nullptr;
  return r;
}

// ------------------------------------------------------------ index writer
IndexFileOut::IndexFileOut(const std::string &path, int mode, bool headers) {
  static const char *ext[3] = {".txt", ".bin", ".bool"};
  static const char *fmt[3] = {"ascii", "binary", "bool"};
  static const char *ver[3] = {"v1.0", "v1.0", "v2.0"};
  if (mode < 0 || mode > 2) throw std::runtime_error("Invalid value passed by user for argument print_mode");
  f_ = fopen((path + ext[mode]).c_str(), "wb");
  if (!f_) throw std::runtime_error(output_open_error(path + ext[mode]));
  if (headers) {
    write_sized_string(f_, fmt[mode]);
    write_sized_string(f_, ver[mode]);
  }
}

IndexFileOut::~IndexFileOut() { fclose(f_); }

void IndexFileOut::write(const char *p, u64 n) { write_all(f_, p, n); }

// ------------------------------------------------------------ color output
ColorFileOut::ColorFileOut(const std::string &path, int mode, bool headers, u64 num_colors) {
  static const char *ext[3] = {".txt", ".bin", ".csv"};
  if (mode < 0 || mode > 2) throw std::runtime_error("Invalid value passed by user for argument print_mode");
  f_ = fopen((path + ext[mode]).c_str(), "wb");
  if (!f_) throw std::runtime_error(output_open_error(path + ext[mode]));
  if (mode == 2) {
    std::string h;
    for (u64 i = 1; i <= num_colors; ++i) {
      h += std::to_string(i);
      h += i == num_colors ? '\n' : ',';
    }
    write_all(f_, h.data(), h.size());
  } else if (headers) {
    write_sized_string(f_, mode == 0 ? "ascii" : "binary");
    write_sized_string(f_, "v1.0");
  }
}

ColorFileOut::~ColorFileOut() { fclose(f_); }

void ColorFileOut::write(const char *p, u64 n) { write_all(f_, p, n); }

// ------------------------------------------------------ index file reading
IndexFileIn::IndexFileIn(const std::string &path) {
  f_ = fopen(path.c_str(), "rb");
  if (!f_) throw InputMissing("The input file " + path + " cannot be opened");
  auto sized = [&]() {
    u64 n = 0;
    if (fread(&n, 8, 1, f_) != 1 || n > 64) {
      fclose(f_);
      throw InputMissing("Invalid file format in file: " + path);
    }
    std::string s(n, '\0');
    if (n && fread(&s[0], 1, n, f_) != n) {
      fclose(f_);
      throw InputMissing("Invalid file format in file: " + path);
    }
    return s;
  };
  const std::string fmt = sized();
  if (fmt == "ascii") format_ = 0;
  else if (fmt == "binary") format_ = 1;
  else {
    fclose(f_);
    throw InputMissing("Invalid file format in file: " + path);
  }
  if (sized() != "v1.0") {
    fclose(f_);
    throw std::runtime_error("The file has an incompatible version number");
  }
}

IndexFileIn::~IndexFileIn() { fclose(f_); }

u64 IndexFileIn::read(char *p, u64 n) {
  u64 got = 0;
  while (got < n) {
    const size_t m = fread(p + got, 1, n - got, f_);
    if (m == 0) {
      if (ferror(f_)) throw std::runtime_error("read failed");
      break;
    }
    got += m;
  }
  return got;
}

std::vector<std::string> expand_list(const std::string &name) {
  const std::string ext = ".list";
  if (name.size() >= ext.size() && name.compare(name.size() - ext.size(), ext.size(), ext) == 0) {
    std::ifstream in(name);
    if (!in) throw std::runtime_error("The input file " + name + " cannot be opened");
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
