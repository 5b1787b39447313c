// Device-side formatting of index results into the reference's output bytes
// (src/ResultsPrinter/AsciiContinuousIndexResultsPrinter.cpp:50-85,
// BinaryContinuousIndexResultsPrinter.cpp:52-78,
// BoolContinuousIndexResultsPrinter.cpp:43-69; headers are the caller's):
//   ascii   results of a read separated by ' ', "-1" not found, "-2" invalid,
//           '\n' at the end of every read (an empty line for a read < k)
//   binary  one u64 per result, u64 max - 2 at the end of every read
//   bool    '0' found, '1' not found, '2' invalid, '\n' at the end of a read
// A batch is a list of read pieces: piece e holds the k-mers
// [kmer_ends[e-1], kmer_ends[e]) and closes its read when closes[e] != 0;
// otherwise the next piece (or the next batch, line_open) continues its line.
#pragma once

#include "common.hpp"

namespace sbwt_amd {

constexpr int kFormatAscii = 0, kFormatBinary = 1, kFormatBool = 2;

class Formatter {
 public:
  Formatter(u64 max_kmers, u64 max_pieces, hipStream_t stream);
  ~Formatter();
  // Bytes of the formatted batch; written to d_out when d_out is not null
  // and capacity suffices. Synchronous (the byte count comes back to the host).
  // The k-mer count is the last piece end (read back from the device).
  u64 format(int mode, const u64 *d_results, const u64 *d_piece_kmer_ends, const u8 *d_piece_closes,
             u64 n_pieces, bool line_open, char *d_out, u64 capacity);
  hipStream_t stream() const { return stream_; }
  // Host read-closing flags in, host bytes out: the batch is formatted into
  // the formatter's own device buffer (grown as needed) and copied to
  // host_out when capacity suffices; copy_last copies it later otherwise.
  u64 format_to_host(int mode, const u64 *d_results, const u64 *d_piece_kmer_ends,
                     const u8 *host_closes, u64 n_pieces, bool line_open, char *host_out,
                     u64 capacity);
  void copy_last(char *host_out, u64 capacity);

 private:
  u64 max_items_;
  hipStream_t stream_;
  bool own_stream_;
  u64 *d_block_sums_ = nullptr;  // per 256 items, then their inclusive scan
  u64 *d_scan_tmp_ = nullptr;
  u8 *d_closes_ = nullptr;   // staged closing flags (max_pieces)
  u64 max_pieces_ = 0;
  char *d_out_ = nullptr;    // formatted bytes of the last format_to_host call
  u64 out_capacity_ = 0, last_bytes_ = 0;
};

}  // namespace sbwt_amd
