// Device parser of index-result files, the input of the colors stage
// (src/IndexFileParser/AsciiIndexFileParser.cpp:33-98,
// BinaryIndexFileParser.cpp:33-70, IndexFileParser.cpp:36-74).
//
// A body (the file after its two sized header strings) becomes
//   results[i]    found index, kNotFound for a counted "-1" / u64 max,
//                 kInvalid for a counted "-2" / u64 max - 1, in file order
//   kmer_ends[r]  number of results before the end of read r
// which is the per-k-mer form the color search takes. Results after the last
// read end belong to no read.
//
// Ascii follows the reference's char loop exactly: chars are signed and any
// char >= '0' is a digit; a digit run is a number (u64 arithmetic, wrapping);
// '-' consumes the next char (whatever it is: '1' counts a not-found, '2' an
// invalid), then every digit after it, then one more char, which ends the
// read only if it is '\n'; a number's terminator is read again as the start
// of what follows; every other '\n' ends a read. On the device this loop is a
// four-state automaton: each 16-byte segment is summarised by its state map
// (start state -> end state), the maps are scanned across the body, and each
// segment is then walked from its known start state. A batch boundary is
// safe right after any '\n' whose preceding byte is not '-' (the automaton is
// then back in its top state): batch_cut() finds one.
// Binary is one u64 word per entry: u64 max - 2 ends a read, anything else is
// a result; a partial last word is not read.
#pragma once

#include "common.hpp"

namespace sbwt_amd {

constexpr int kIndexFileAscii = 0, kIndexFileBinary = 1;

class IndexFileParser {
 public:
  // max_bytes < 2^32 per batch; at most max_reads read ends per batch.
  IndexFileParser(u64 max_bytes, u64 max_reads, hipStream_t stream);
  ~IndexFileParser();
  IndexFileParser(const IndexFileParser &) = delete;
  IndexFileParser &operator=(const IndexFileParser &) = delete;

  // Parses n_bytes of a body at d_body (device, 16-byte aligned) into
  // results() / kmer_ends(). Returns the number of reads; *n_results the
  // number of results. Throws when the batch holds more than max_reads read
  // ends. Synchronous on the counts; the entries are complete when the
  // stream reaches them.
  u64 parse_device(int format, const char *d_body, u64 n_bytes, u64 *n_results);
  // The same from host bytes (staged into the parser's device buffer).
  u64 parse_host(int format, const char *body, u64 n_bytes, u64 *n_results);

  const u64 *results() const { return d_results_; }
  const u64 *kmer_ends() const { return d_kends_; }
  u64 max_results() const { return max_results_; }
  u64 max_reads() const { return max_reads_; }
  hipStream_t stream() const { return stream_; }

  // Largest prefix length of body[0, n) that ends right after a read end at
  // which the ascii automaton is in its top state (a '\n' not preceded by
  // '-'), or after a u64 max - 2 word in binary, holding at most max_reads
  // read ends; 0 if there is none.
  static u64 batch_cut(int format, const char *body, u64 n, u64 max_reads);
  // Longest prefix of body[0, n) (which starts at a read start) ending right
  // after the read end of a read with a found index, or 0. A batch cut there
  // keeps every read without a found index in the batch of the next read that
  // has one, whose counts the reference prints it with (colored_seq_id,
  // IndexFileParser.cpp:65-74). Read ends as in batch_cut; a read's tokens
  // follow the reference's char loop (AsciiIndexFileParser.cpp:49-66).
  static u64 colored_cut(int format, const char *body, u64 n);

 private:
  u64 max_bytes_, max_reads_, max_results_;
  u64 block_cap_ = 0;             // entries per block array (binary blocks: 2 KiB)
  hipStream_t stream_;
  bool own_stream_;
  char *d_body_ = nullptr;        // staging for host bodies
  u64 *d_results_ = nullptr;
  u64 *d_kends_ = nullptr;
  u8 *d_block_maps_ = nullptr;    // per block: state map, then start state
  u64 *d_block_counts_ = nullptr; // per block: results << 32 | read ends, then the scan
  u64 *d_scan_tmp_ = nullptr;
};

}  // namespace sbwt_amd
