// Device parser of FASTA chunks for the index command: raw file bytes that
// begin at a record's header line and end at a record's end become the
// concatenated sequence chars and the cumulative end of every read, the
// input of the fused search. The record grammar is the host reader's (kseq's,
// src/SequenceFileParser/ContinuousSequenceFileParser.cpp:125,152) for files
// with no line starting with '+': a line starting with '>' or '@' is a
// header, every other line (an empty one included) is sequence, whose bytes
// other than '\n' are chars ('\r' included); a last line without '\n' counts.
// A line starting with '+' (a FASTQ quality block) is out of this grammar:
// the caller keeps such input on the host reader, and the parser rejects it.
#pragma once

#include "common.hpp"

namespace sbwt_amd {

class FastaParser {
 public:
  // At most max_bytes (< 2^32) per chunk and max_reads records per chunk.
  FastaParser(u64 max_bytes, u64 max_reads, hipStream_t stream);
  ~FastaParser();
  FastaParser(const FastaParser &) = delete;
  FastaParser &operator=(const FastaParser &) = delete;

  // Parses host bytes (copied to the device). Returns the number of reads;
  // *n_chars the sequence chars. Throws when the chunk does not start with a
  // header, holds a line starting with '+', or more than max_reads records.
  // Synchronous on the counts.
  u64 parse_host(const char *raw, u64 n_bytes, u64 *n_chars);

  const char *chars() const { return d_chars_; }
  const u64 *read_ends() const { return d_ends_; }
  const u8 *ones() const { return d_ones_; }  // max_reads bytes of 1 (every read closes)
  hipStream_t stream() const { return stream_; }

  // Longest prefix of raw[0, n) that ends right before a header line (or
  // the whole of it at the end of the input, `last`), holding at most
  // max_reads records; 0 if there is none (one record is larger than n).
  // Host only.
  static u64 chunk_cut(const char *raw, u64 n, bool last, u64 max_reads);
  // Whether raw[0, n) holds a line starting with '+' (the first line
  // included when at_line_start). Host only.
  static bool has_quality_line(const char *raw, u64 n, bool at_line_start);

 private:
  u64 max_bytes_, max_reads_;
  hipStream_t stream_;
  bool own_stream_;
  u64 block_cap_ = 0;
  char *d_raw_ = nullptr;
  char *d_chars_ = nullptr;
  u64 *d_ends_ = nullptr;
  u8 *d_ones_ = nullptr;
  u8 *d_block_maps_ = nullptr;     // per block: state map, then start state
  u64 *d_block_counts_ = nullptr;  // per block: chars << 32 | headers, then the scan
  u64 *d_scan_tmp_ = nullptr;
  u32 *d_flags_ = nullptr;         // a '+' line start was seen
};

// Device parser of FASTQ chunks in four-line records (header, one sequence
// line, '+' line, one quality line), the common form of sequencer output.
// The host reader (kseq's rules, as above) reads such a record as its
// sequence line exactly when the header line starts with '@' or '>', the
// sequence line is not empty and does not start with '+', '@' or '>', the
// third line starts with '+', and the quality line is at least as long as
// the sequence line (it then takes that one line as the whole quality and
// the next record starts right after it). The parser checks this per record
// and parses the longest prefix of whole records that passes; the caller
// gives what follows to the host reader when the prefix stops short.
class FastqParser {
 public:
  FastqParser(u64 max_bytes, u64 max_reads, hipStream_t stream);
  ~FastqParser();
  FastqParser(const FastqParser &) = delete;
  FastqParser &operator=(const FastqParser &) = delete;

  // raw[0, n) starts at a record's header line; `last` when it ends the file
  // (a last line without '\n' then counts). Parses the longest prefix of at
  // most max_reads whole four-line records that read as above and returns its
  // length in bytes (0: no such record); *n_reads and *n_chars describe it.
  // Synchronous.
  u64 parse_host(const char *raw, u64 n, bool last, u64 *n_reads, u64 *n_chars);

  const char *chars() const { return d_chars_; }
  const u64 *read_ends() const { return d_ends_; }
  const u8 *ones() const { return d_ones_; }
  hipStream_t stream() const { return stream_; }
  u64 max_bytes() const { return max_bytes_; }
  u64 max_reads() const { return max_reads_; }

 private:
  u64 max_bytes_, max_reads_;
  hipStream_t stream_;
  bool own_stream_;
  u64 block_cap_ = 0;
  char *d_raw_ = nullptr;
  char *d_chars_ = nullptr;
  u64 *d_ends_ = nullptr;
  u8 *d_ones_ = nullptr;
  u32 *d_line_ends_ = nullptr;     // byte offset of each line's '\n' (n for a last line without)
  u8 *d_block_maps_ = nullptr;     // per block: state map, then start state
  u64 *d_block_counts_ = nullptr;  // per block: sequence chars << 32 | newlines, then the scan
  u64 *d_scan_tmp_ = nullptr;
  u32 *d_first_bad_ = nullptr;     // first record that fails the check
  u64 *d_result_ = nullptr;        // records parsed, prefix bytes, chars
};

}  // namespace sbwt_amd
