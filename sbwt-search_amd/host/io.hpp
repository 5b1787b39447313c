// Host I/O of the sbwt_search command line: sequence files, index-result files
// and the output formats of the reference's printers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sbwt_host {

using u32 = uint32_t;
using u64 = uint64_t;

// An input file that cannot be read as such: the reference logs it and goes
// on with the next file, whose output file is still written (with headers
// only) (ContinuousSequenceFileParser.cpp:113-132,
// ContinuousIndexFileParser.cpp:64-100).
struct InputMissing : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// FASTA/FASTQ reader (optionally gzipped). The reference delegates to
// kseqpp_REad v1.6.0 (src/SequenceFileParser/ContinuousSequenceFileParser.cpp
// :125,152); the record grammar is kseq's: skip to the first '>' or '@', the
// header is the rest of that line, the sequence is the concatenation of the
// following lines up to a line starting with '>', '@' or '+'; after '+' the
// quality lines are consumed until they are as long as the sequence.
class SeqReader {
 public:
  explicit SeqReader(const std::string &path);
  // Continues an open file (a gzFile, now owned by the reader) whose next
  // bytes are pending[0, n) and then the rest of the file.
  SeqReader(void *gz, const char *pending, u64 n);
  ~SeqReader();
  SeqReader(const SeqReader &) = delete;
  SeqReader &operator=(const SeqReader &) = delete;
  // Appends the next record's sequence to `chars`; false at end of file.
  bool next(std::vector<char> &chars);

 private:
  bool fill();
  int getc();
  int peek();
  u64 take_line(std::vector<char> *out);
  void *gz_ = nullptr;
  std::vector<char> buf_;
  size_t pos_ = 0, len_ = 0;
  bool eof_ = false;
  bool started_ = false;
  int pending_ = -1;  // first char of the next header, already read
};

// A sequence file read as raw (decompressed) bytes, the input of the device
// FASTA parser. to_reader() hands the open file to a SeqReader that goes on
// from bytes already read, when the rest is for the host reader.
class RawSeqFile {
 public:
  explicit RawSeqFile(const std::string &path);  // InputMissing when it cannot be opened
  ~RawSeqFile();
  RawSeqFile(const RawSeqFile &) = delete;
  RawSeqFile &operator=(const RawSeqFile &) = delete;
  // Reads up to n bytes into dst; fewer only at the end of the file.
  u64 read(char *dst, u64 n);
  bool at_end() const { return eof_; }
  std::unique_ptr<SeqReader> to_reader(const char *pending, u64 n);

 private:
  void *gz_ = nullptr;
  bool eof_ = false;
};

// Index output file (src/IndexResultsPrinter/*): the header (u64 length +
// "ascii"/"binary"/"bool", u64 length + "v1.0"/"v1.0"/"v2.0") unless
// headers are off, then the formatted bytes (formatted on the device).
class IndexFileOut {
 public:
  IndexFileOut(const std::string &path, int mode, bool headers);
  ~IndexFileOut();
  void write(const char *p, u64 n);

 private:
  FILE *f_;
};

// Color output file (src/ColorResultsPrinter/*): ascii and binary start with
// the header (u64 length + "ascii"/"binary", u64 length + "v1.0") unless
// headers are off; csv always starts with the row "1,2,...,num_colors"
// (CsvContinuousColorResultsPrinter.cpp:60-75). Then the printer bytes
// (formatted on the device). mode: 0 ascii (.txt), 1 binary (.bin), 2 csv (.csv).
class ColorFileOut {
 public:
  ColorFileOut(const std::string &path, int mode, bool headers, u64 num_colors);
  ~ColorFileOut();
  void write(const char *p, u64 n);

 private:
  FILE *f_;
};

// Index-result file opened for the colors stage: its two sized header strings
// are read and checked (ContinuousIndexFileParser.cpp:84-99,
// AsciiIndexFileParser.cpp:25-30, BinaryIndexFileParser.cpp:25-30); read()
// then returns body bytes.
class IndexFileIn {
 public:
  explicit IndexFileIn(const std::string &path);
  ~IndexFileIn();
  int format() const { return format_; }  // 0 ascii, 1 binary
  // Up to n body bytes into p; fewer only at the end of the file.
  u64 read(char *p, u64 n);

 private:
  FILE *f_;
  int format_ = 0;
};

void write_sized_string(FILE *f, const std::string &s);
std::vector<std::string> expand_list(const std::string &name);
u64 file_size(const std::string &path);
// LPT partition of files into streams by size, largest first
// (src/FilesizeLoadBalancer/FilesizeLoadBalancer.cpp:23-55).
std::vector<std::vector<u64>> partition_by_size(const std::vector<std::string> &files, u64 parts);

}  // namespace sbwt_host
