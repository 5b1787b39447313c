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
  // The same with the closing flags already on the device.
  u64 format_device_closes_to_host(int mode, const u64 *d_results, const u64 *d_piece_kmer_ends,
                                   const u8 *d_closes, u64 n_pieces, bool line_open,
                                   char *host_out, u64 capacity);
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

// Color id lists of a batch of reads (the color stage's CSR: d_offsets[r] ..
// d_offsets[r+1] into d_ids, ascending per read) as the reference's color
// printers write them, without file headers
// (src/ColorResultsPrinter/AsciiContinuousColorResultsPrinter.cpp:62-79,
// BinaryContinuousColorResultsPrinter.cpp:51-62,
// CsvContinuousColorResultsPrinter.cpp:40-50,78-110, line structure
// ContinuousColorResultsPrinter.hpp:237-263):
//   ascii   ids separated by ' ', '\n' at the end of every read
//   binary  one u64 per id, u64 max at the end of every read
//   csv     per read num_colors "0"/"1" joined by ',' and ended by '\n'
constexpr int kColorFormatAscii = 0, kColorFormatBinary = 1, kColorFormatCsv = 2;

class ColorFormatter {
 public:
  ColorFormatter(u64 max_ids, u64 max_reads, hipStream_t stream);
  ~ColorFormatter();
  ColorFormatter(const ColorFormatter &) = delete;
  ColorFormatter &operator=(const ColorFormatter &) = delete;
  // Bytes of the batch; written to d_out when not null and capacity
  // suffices. Synchronous on the byte count.
  u64 format(int mode, const u64 *d_offsets, const u32 *d_ids, u64 n_reads, u64 num_colors,
             char *d_out, u64 capacity);
  // Into the formatter's own device buffer (grown as needed), then to
  // host_out when capacity suffices; copy_last copies it later otherwise.
  u64 format_to_host(int mode, const u64 *d_offsets, const u32 *d_ids, u64 n_reads,
                     u64 num_colors, char *host_out, u64 capacity);
  void copy_last(char *host_out, u64 capacity);
  hipStream_t stream() const { return stream_; }

 private:
  u64 max_items_, max_reads_;
  hipStream_t stream_;
  bool own_stream_;
  u64 *d_block_sums_ = nullptr;
  u64 *d_scan_tmp_ = nullptr;
  char *d_out_ = nullptr;
  u64 out_capacity_ = 0, last_bytes_ = 0;
};

}  // namespace sbwt_amd
