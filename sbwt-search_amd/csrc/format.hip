// Device-side formatting of index results (see format.hpp). Two passes over
// "items": every k-mer of a piece, then one end item per piece (the read's
// newline, or nothing when the piece does not close its read). Pass one
// sums the byte lengths of each 256-item block; a device scan turns the sums
// into block offsets; pass two recomputes the lengths, takes the block-local
// exclusive prefix and writes each item's bytes at its offset.
#include "format.hpp"

#include <algorithm>
#include <stdexcept>

#include "scan.hpp"

namespace sbwt_amd {
namespace {

constexpr u32 kFmtThreads = 256;

struct Item {
  bool kmer;     // a k-mer result (else the piece's end item)
  bool sep;      // ascii: a ' ' goes before the result
  bool closes;   // end item: the piece ends its read
  u64 value;     // k-mer result
};

__device__ __forceinline__ u32 dec_digits(u64 v) {
  u32 d = 1;
  u64 p = 10;
  while (d < 20 && v >= p) {
    ++d;
    p *= 10;
  }
  return d;
}

template <int MODE>
__device__ __forceinline__ u32 item_len(const Item &it) {
  if (!it.kmer) {
    if (!it.closes) return 0;
    return MODE == kFormatBinary ? 8u : 1u;
  }
  if constexpr (MODE == kFormatBinary) return 8;
  if constexpr (MODE == kFormatBool) return 1;
  const u32 text = (it.value == kInvalid || it.value == kNotFound) ? 2u : dec_digits(it.value);
  return text + (it.sep ? 1u : 0u);
}

template <int MODE, u64 END_WORD>
__device__ __forceinline__ void item_write(const Item &it, char *o) {
  if (!it.kmer) {
    if (!it.closes) return;
    if constexpr (MODE == kFormatBinary) {
      const u64 nl = END_WORD;
#pragma unroll
      for (u32 i = 0; i < 8; ++i) o[i] = static_cast<char>((nl >> (8 * i)) & 0xff);
    } else {
      o[0] = '\n';
    }
    return;
  }
  const u64 v = it.value;
  if constexpr (MODE == kFormatBinary) {
#pragma unroll
    for (u32 i = 0; i < 8; ++i) o[i] = static_cast<char>((v >> (8 * i)) & 0xff);
  } else if constexpr (MODE == kFormatBool) {
    o[0] = v == kInvalid ? '2' : (v == kNotFound ? '1' : '0');
  } else {
    if (it.sep) *o++ = ' ';
    if (v == kInvalid || v == kNotFound) {
      o[0] = '-';
      o[1] = v == kInvalid ? '2' : '1';
      return;
    }
    const u32 d = dec_digits(v);
    if (v < (1ULL << 32)) {
      u32 x = static_cast<u32>(v);
      for (u32 i = d; i-- > 0;) {
        o[i] = static_cast<char>('0' + x % 10u);
        x /= 10u;
      }
    } else {
      u64 x = v;
      for (u32 i = d; i-- > 0;) {
        o[i] = static_cast<char>('0' + x % 10u);
        x /= 10u;
      }
    }
  }
}

// First piece e whose items end after t: items of piece e are
// [kends[e-1] + e, kends[e] + e] (k-mers, then the end item).
__device__ __forceinline__ u64 piece_of(const u64 *kends, u64 lo, u64 hi, u64 t) {
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (kends[mid] + mid + 1 > t) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Index results: pieces of reads, sentinels printed as "-1" / "-2", binary
// read end u64 max - 2.
struct IndexItems {
  const u64 *results;
  const u64 *kends;
  const u8 *closes;
  u64 n_pieces;
  u64 n_kmers;
  u32 line_open;
  static constexpr u64 kEndWord = kNewline;

  // The item t of a block whose pieces lie in [lo, hi] (found by block_bounds).
  __device__ __forceinline__ Item at(u64 t, u64 n_items, u64 lo, u64 hi) const {
    Item it{false, false, false, 0};
    if (t >= n_items) return it;
    const u64 e = piece_of(kends, lo, min(hi + 1, n_pieces), t);
    const u64 ks = e ? kends[e - 1] : 0;
    const u64 nk = kends[e] > ks ? kends[e] - ks : 0;  // piece ends must not decrease
    const u64 j = t - (ks + e);
    if (j < nk) {
      it.kmer = true;
      it.value = ks + j < n_kmers ? results[ks + j] : kNotFound;
      // a ' ' goes before every result but the first of its line: look back
      // over earlier pieces of the same line for one with results
      bool sep = j > 0;
      if (!sep) {
        sep = line_open != 0;
        for (u64 q = e; q-- > 0;) {
          if (closes[q]) {
            sep = false;
            break;
          }
          if (kends[q] > (q ? kends[q - 1] : 0)) {
            sep = true;
            break;
          }
        }
      }
      it.sep = sep;
    } else {
      it.closes = closes[e] != 0;
    }
    return it;
  }
};

// Color id lists: one line per read (ids[offsets[r] .. offsets[r+1])), binary
// read end u64 max; kends = offsets + 1.
struct ColorItems {
  const u32 *ids;
  const u64 *kends;
  u64 n_pieces;
  static constexpr u64 kEndWord = ~0ULL;

  __device__ __forceinline__ Item at(u64 t, u64 n_items, u64 lo, u64 hi) const {
    Item it{false, false, false, 0};
    if (t >= n_items) return it;
    const u64 e = piece_of(kends, lo, min(hi + 1, n_pieces), t);
    const u64 ks = e ? kends[e - 1] : 0;
    const u64 j = t - (ks + e);
    if (j < kends[e] - ks) {
      it.kmer = true;
      it.value = ids[ks + j];
      it.sep = j > 0;
    } else {
      it.closes = true;
    }
    return it;
  }
};

__device__ __forceinline__ void block_bounds(u64 blk, u64 n_items, const u64 *kends, u64 n_pieces,
                                             u64 *sh) {
  const u32 tid = threadIdx.x;
  if (tid == 0 || tid == kFmtThreads - 1) {
    const u64 t = min(blk * kFmtThreads + tid, n_items - 1);
    sh[tid == 0 ? 0 : 1] = piece_of(kends, 0, n_pieces, t);
  }
  __syncthreads();
}

// Block-wide sum (all threads get it) and exclusive prefix of x.
__device__ __forceinline__ u32 block_prefix(u32 x, u32 *sh_waves, u32 &total) {
  const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  u32 incl = x;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) sh_waves[wave] = incl;
  __syncthreads();
  u32 before = 0;
  total = 0;
#pragma unroll
  for (u32 w = 0; w < kFmtThreads / 64; ++w) {
    if (w < wave) before += sh_waves[w];
    total += sh_waves[w];
  }
  __syncthreads();
  return before + incl - x;
}

template <int MODE, typename Src>
__global__ __launch_bounds__(kFmtThreads) void k_fmt_count(Src src, u64 n_items,
                                                           u64 *__restrict__ block_sums) {
  __shared__ u64 sh_bounds[2];
  __shared__ u32 sh_waves[kFmtThreads / 64];
  const u64 n_blocks = (n_items + kFmtThreads - 1) / kFmtThreads;
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    block_bounds(blk, n_items, src.kends, src.n_pieces, sh_bounds);
    const Item it = src.at(blk * kFmtThreads + threadIdx.x, n_items, sh_bounds[0], sh_bounds[1]);
    u32 total = 0;
    block_prefix(item_len<MODE>(it), sh_waves, total);
    if (threadIdx.x == 0) block_sums[blk] = total;
  }
}

template <int MODE, typename Src>
__global__ __launch_bounds__(kFmtThreads) void k_fmt_write(Src src, u64 n_items,
                                                           const u64 *__restrict__ block_incl,
                                                           char *__restrict__ out) {
  __shared__ u64 sh_bounds[2];
  __shared__ u32 sh_waves[kFmtThreads / 64];
  const u64 n_blocks = (n_items + kFmtThreads - 1) / kFmtThreads;
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    block_bounds(blk, n_items, src.kends, src.n_pieces, sh_bounds);
    const Item it = src.at(blk * kFmtThreads + threadIdx.x, n_items, sh_bounds[0], sh_bounds[1]);
    u32 total = 0;
    const u32 off = block_prefix(item_len<MODE>(it), sh_waves, total);
    const u64 base = blk ? block_incl[blk - 1] : 0;
    item_write<MODE, Src::kEndWord>(it, out + base + off);
  }
}

// CSV rows (CsvContinuousColorResultsPrinter.cpp:40-50,78-110): every row is
// num_colors "0" joined by ',' and ended by '\n'; then '1' at 2 * id.
__global__ __launch_bounds__(kFmtThreads) void k_csv_fill(char *__restrict__ out, u64 n_bytes,
                                                          u64 row, u32 aligned) {
  const u64 chunks = (n_bytes + 15) / 16;
  for (u64 c = blockIdx.x * u64{kFmtThreads} + threadIdx.x; c < chunks;
       c += u64{gridDim.x} * kFmtThreads) {
    const u64 beg = c * 16;
    u64 q = beg % row;
    u32 w[4] = {0, 0, 0, 0};
#pragma unroll
    for (u32 i = 0; i < 16; ++i) {
      const u32 ch = (q & 1) ? (q + 1 == row ? '\n' : ',') : '0';
      w[i / 4] |= ch << (8 * (i % 4));
      q = q + 1 == row ? 0 : q + 1;
    }
    if (aligned && beg + 16 <= n_bytes) {
      *reinterpret_cast<uint4 *>(out + beg) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
      for (u32 i = 0; i < 16 && beg + i < n_bytes; ++i) out[beg + i] = static_cast<char>(w[i / 4] >> (8 * (i % 4)));
    }
  }
}

__global__ __launch_bounds__(kFmtThreads) void k_csv_ones(const u64 *__restrict__ offsets,
                                                          const u32 *__restrict__ ids, u64 n_reads,
                                                          u64 row, char *__restrict__ out) {
  for (u64 r = blockIdx.x * u64{kFmtThreads} + threadIdx.x; r < n_reads;
       r += u64{gridDim.x} * kFmtThreads) {
    for (u64 i = offsets[r]; i < offsets[r + 1]; ++i) {
      const u64 c = ids[i];
      if (2 * c < row) out[r * row + 2 * c] = '1';  // ids past num_colors have no column
    }
  }
}

template <typename F>
void dispatch_mode(int mode, F f) {
  if (mode == kFormatAscii) f.template operator()<kFormatAscii>();
  else if (mode == kFormatBinary) f.template operator()<kFormatBinary>();
  else if (mode == kFormatBool) f.template operator()<kFormatBool>();
  else throw std::runtime_error("unknown output format");
}

}  // namespace

Formatter::Formatter(u64 max_kmers, u64 max_pieces, hipStream_t stream)
    : max_items_(max_kmers + max_pieces), stream_(stream), own_stream_(stream == nullptr),
      max_pieces_(max_pieces) {
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  const u64 blocks = max_items_ / kFmtThreads + 2;
  SBWT_HIP(hipMalloc(&d_block_sums_, blocks * 2 * 8));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(blocks) * 8));
}

Formatter::~Formatter() {
  if (d_block_sums_) (void)hipFree(d_block_sums_);
  if (d_scan_tmp_) (void)hipFree(d_scan_tmp_);
  if (d_closes_) (void)hipFree(d_closes_);
  if (d_out_) (void)hipFree(d_out_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

u64 Formatter::format(int mode, const u64 *d_results, const u64 *d_piece_kmer_ends,
                      const u8 *d_piece_closes, u64 n_pieces, bool line_open, char *d_out,
                      u64 capacity) {
  if (n_pieces == 0) return 0;
  u64 n_kmers = 0;
  SBWT_HIP(hipMemcpyAsync(&n_kmers, d_piece_kmer_ends + n_pieces - 1, 8, hipMemcpyDeviceToHost,
                          stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  const u64 n_items = n_kmers + n_pieces;
  if (n_items > max_items_)
    throw std::runtime_error("batch exceeds the formatter's max_kmers / max_pieces");
  const u64 n_blocks = (n_items + kFmtThreads - 1) / kFmtThreads;
  const u32 grid = static_cast<u32>(std::min<u64>(n_blocks, 1u << 20));
  u64 *incl = d_block_sums_ + (max_items_ / kFmtThreads + 2);
  const IndexItems src{d_results, d_piece_kmer_ends, d_piece_closes, n_pieces, n_kmers,
                       line_open ? 1u : 0u};
  dispatch_mode(mode, [&]<int MODE>() {
    k_fmt_count<MODE><<<grid, kFmtThreads, 0, stream_>>>(src, n_items, d_block_sums_);
  });
  SBWT_HIP(hipGetLastError());
  scan_inclusive_u64(d_block_sums_, n_blocks, incl, d_scan_tmp_, stream_);
  u64 bytes = 0;
  SBWT_HIP(hipMemcpyAsync(&bytes, incl + n_blocks - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  if (d_out && capacity >= bytes && bytes) {
    dispatch_mode(mode, [&]<int MODE>() {
      k_fmt_write<MODE><<<grid, kFmtThreads, 0, stream_>>>(src, n_items, incl, d_out);
    });
    SBWT_HIP(hipGetLastError());
  }
  return bytes;
}

u64 Formatter::format_to_host(int mode, const u64 *d_results, const u64 *d_piece_kmer_ends,
                              const u8 *host_closes, u64 n_pieces, bool line_open, char *host_out,
                              u64 capacity) {
  last_bytes_ = 0;
  if (n_pieces == 0) return 0;
  if (n_pieces > max_pieces_) throw std::runtime_error("batch exceeds the formatter's max_pieces");
  if (!d_closes_) SBWT_HIP(hipMalloc(&d_closes_, max_pieces_));
  SBWT_HIP(hipMemcpyAsync(d_closes_, host_closes, n_pieces, hipMemcpyHostToDevice, stream_));
  return format_device_closes_to_host(mode, d_results, d_piece_kmer_ends, d_closes_, n_pieces,
                                      line_open, host_out, capacity);
}

u64 Formatter::format_device_closes_to_host(int mode, const u64 *d_results,
                                            const u64 *d_piece_kmer_ends, const u8 *d_closes,
                                            u64 n_pieces, bool line_open, char *host_out,
                                            u64 capacity) {
  last_bytes_ = 0;
  if (n_pieces == 0) return 0;
  // written at once when the device buffer holds it; else grown and formatted again
  const u64 bytes = format(mode, d_results, d_piece_kmer_ends, d_closes, n_pieces, line_open,
                           d_out_, out_capacity_);
  if (bytes > out_capacity_) {
    if (d_out_) SBWT_HIP(hipFree(d_out_));
    d_out_ = nullptr;
    out_capacity_ = 0;
    const u64 cap = bytes + bytes / 4 + 4096;
    SBWT_HIP(hipMalloc(&d_out_, cap));
    out_capacity_ = cap;
    format(mode, d_results, d_piece_kmer_ends, d_closes, n_pieces, line_open, d_out_, out_capacity_);
  }
  last_bytes_ = bytes;
  if (host_out && capacity >= bytes) copy_last(host_out, capacity);
  return bytes;
}

void Formatter::copy_last(char *host_out, u64 capacity) {
  if (capacity < last_bytes_) throw std::runtime_error("output buffer too small for the batch");
  if (last_bytes_)
    SBWT_HIP(hipMemcpyAsync(host_out, d_out_, last_bytes_, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

}  // namespace sbwt_amd

namespace sbwt_amd {

ColorFormatter::ColorFormatter(u64 max_ids, u64 max_reads, hipStream_t stream)
    : max_items_(max_ids + max_reads), max_reads_(max_reads), stream_(stream),
      own_stream_(stream == nullptr) {
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  const u64 blocks = max_items_ / kFmtThreads + 2;
  SBWT_HIP(hipMalloc(&d_block_sums_, blocks * 2 * 8));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(blocks) * 8));
}

ColorFormatter::~ColorFormatter() {
  if (d_block_sums_) (void)hipFree(d_block_sums_);
  if (d_scan_tmp_) (void)hipFree(d_scan_tmp_);
  if (d_out_) (void)hipFree(d_out_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

u64 ColorFormatter::format(int mode, const u64 *d_offsets, const u32 *d_ids, u64 n_reads,
                           u64 num_colors, char *d_out, u64 capacity) {
  if (mode != kColorFormatAscii && mode != kColorFormatBinary && mode != kColorFormatCsv)
    throw std::runtime_error("unknown color output format");
  if (n_reads == 0) return 0;
  if (n_reads > max_reads_) throw std::runtime_error("batch exceeds the color formatter's max reads");
  const u32 grid_cap = 1u << 16;
  if (mode == kColorFormatCsv) {
    const u64 row = 2 * num_colors, bytes = n_reads * row;
    if (d_out && capacity >= bytes && bytes) {
      const u64 chunks = (bytes + 15) / 16;
      k_csv_fill<<<static_cast<u32>(std::min<u64>((chunks + kFmtThreads - 1) / kFmtThreads, grid_cap)),
                   kFmtThreads, 0, stream_>>>(d_out, bytes, row,
                                              reinterpret_cast<uintptr_t>(d_out) % 16 == 0 ? 1u : 0u);
      k_csv_ones<<<static_cast<u32>(std::min<u64>((n_reads + kFmtThreads - 1) / kFmtThreads, grid_cap)),
                   kFmtThreads, 0, stream_>>>(d_offsets, d_ids, n_reads, row, d_out);
      SBWT_HIP(hipGetLastError());
    }
    return bytes;
  }
  u64 n_ids = 0;
  SBWT_HIP(hipMemcpyAsync(&n_ids, d_offsets + n_reads, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  const u64 n_items = n_ids + n_reads;
  if (n_items > max_items_) throw std::runtime_error("batch exceeds the color formatter's max ids");
  const u64 n_blocks = (n_items + kFmtThreads - 1) / kFmtThreads;
  const u32 grid = static_cast<u32>(std::min<u64>(n_blocks, 1u << 20));
  u64 *incl = d_block_sums_ + (max_items_ / kFmtThreads + 2);
  const ColorItems src{d_ids, d_offsets + 1, n_reads};
  const int fmt = mode == kColorFormatAscii ? kFormatAscii : kFormatBinary;
  dispatch_mode(fmt, [&]<int MODE>() {
    k_fmt_count<MODE><<<grid, kFmtThreads, 0, stream_>>>(src, n_items, d_block_sums_);
  });
  SBWT_HIP(hipGetLastError());
  scan_inclusive_u64(d_block_sums_, n_blocks, incl, d_scan_tmp_, stream_);
  u64 bytes = 0;
  SBWT_HIP(hipMemcpyAsync(&bytes, incl + n_blocks - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  if (d_out && capacity >= bytes && bytes) {
    dispatch_mode(fmt, [&]<int MODE>() {
      k_fmt_write<MODE><<<grid, kFmtThreads, 0, stream_>>>(src, n_items, incl, d_out);
    });
    SBWT_HIP(hipGetLastError());
  }
  return bytes;
}

u64 ColorFormatter::format_to_host(int mode, const u64 *d_offsets, const u32 *d_ids, u64 n_reads,
                                   u64 num_colors, char *host_out, u64 capacity) {
  last_bytes_ = 0;
  // written at once when the device buffer holds it; else grown and formatted again
  const u64 bytes = format(mode, d_offsets, d_ids, n_reads, num_colors, d_out_, out_capacity_);
  if (bytes > out_capacity_) {
    if (d_out_) SBWT_HIP(hipFree(d_out_));
    d_out_ = nullptr;
    out_capacity_ = 0;
    const u64 cap = bytes + bytes / 4 + 4096;
    SBWT_HIP(hipMalloc(&d_out_, cap));
    out_capacity_ = cap;
    format(mode, d_offsets, d_ids, n_reads, num_colors, d_out_, out_capacity_);
  }
  last_bytes_ = bytes;
  if (host_out && capacity >= bytes) copy_last(host_out, capacity);
  return bytes;
}

void ColorFormatter::copy_last(char *host_out, u64 capacity) {
  if (capacity < last_bytes_) throw std::runtime_error("output buffer too small for the batch");
  if (last_bytes_)
    SBWT_HIP(hipMemcpyAsync(host_out, d_out_, last_bytes_, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
}

}  // namespace sbwt_amd
