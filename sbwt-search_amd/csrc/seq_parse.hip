// Device parser of FASTA chunks (see seq_parse.hpp): a three-state byte
// automaton (line start, header line, sequence line) scanned with the
// kernels of byte_automaton.hpp; the count pass sums chars and header starts
// per block, the write pass copies the chars and writes the read ends.
#include "seq_parse.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "byte_automaton.hpp"
#include "scan.hpp"

namespace sbwt_amd {
namespace {

using automaton::block_prefix_sum;
using automaton::kBlockBytes;
using automaton::kSegBytes;
using automaton::kStatesThreads;
using automaton::load_segment;
using automaton::next_state;
constexpr u32 kThreads = automaton::kThreads;

constexpr u32 kLineStart = 0, kHeader = 1, kSeq = 2;  // state 3 is never entered
constexpr u32 kMapNewline = kLineStart | kLineStart << 2 | kLineStart << 4 | 3u << 6;
constexpr u32 kMapHeaderChar = kHeader | kHeader << 2 | kSeq << 4 | 3u << 6;  // '>' '@'
constexpr u32 kMapOther = kSeq | kHeader << 2 | kSeq << 4 | 3u << 6;

__device__ __forceinline__ bool header_char(u8 b) { return b == '>' || b == '@'; }

struct FastaMap {
  __device__ __forceinline__ u32 operator()(u8 b) const {
    return b == '\n' ? kMapNewline : (header_char(b) ? kMapHeaderChar : kMapOther);
  }
};

// Walks one segment from state st: counts chars (*nc) and header starts
// (*nh); with WRITE also stores the chars at cbase + *nc and, for a header
// start that is not the chunk's first, the end of the read before it.
template <bool WRITE>
__device__ __forceinline__ void walk(const u8 (&b)[kSegBytes], u32 len, u32 st, u32 &nc, u32 &nh,
                                     u32 &plus, char *chars, u64 cbase, u64 *ends, u64 hbase) {
#pragma unroll
  for (u32 i = 0; i < kSegBytes; ++i) {
    if (i < len) {
      const u8 c = b[i];
      const bool header = st == kLineStart && header_char(c);
      const bool ch = c != '\n' && (st == kSeq || (st == kLineStart && !header_char(c)));
      plus |= (st == kLineStart && c == '+') ? 1u : 0u;
      if constexpr (WRITE) {
        if (ch) chars[cbase + nc] = static_cast<char>(c);
        if (header && hbase + nh > 0) ends[hbase + nh - 1] = cbase + nc;
      }
      nc += ch ? 1u : 0u;
      nh += header ? 1u : 0u;
      st = next_state(FastaMap{}(c), st);
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_fasta_maps(const u8 *__restrict__ raw, u64 n,
                                                         u64 n_blocks, u8 *__restrict__ maps) {
  automaton::block_maps(raw, n, n_blocks, maps, FastaMap{});
}

__global__ __launch_bounds__(kStatesThreads) void k_fasta_states(const u8 *__restrict__ maps,
                                                                 u64 n_blocks,
                                                                 u8 *__restrict__ states) {
  automaton::block_states(maps, n_blocks, kLineStart, states);
}

__global__ __launch_bounds__(kThreads) void k_fasta_count(const u8 *__restrict__ raw, u64 n,
                                                          u64 n_blocks,
                                                          const u8 *__restrict__ states,
                                                          u64 *__restrict__ counts,
                                                          u32 *__restrict__ flags) {
  __shared__ u32 sh[kThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(raw, n, blk * kThreads + threadIdx.x, b);
    const u32 st = automaton::segment_start(blk, states, sh, b, len, FastaMap{});
    u32 nc = 0, nh = 0, plus = 0;
    walk<false>(b, len, st, nc, nh, plus, nullptr, 0, nullptr, 0);
    if (plus) atomicOr(flags, 1u);
    u32 tc, th;
    block_prefix_sum(nc, sh, &tc);
    block_prefix_sum(nh, sh, &th);
    if (threadIdx.x == 0) counts[blk] = (u64{tc} << 32) | th;
  }
}

__global__ __launch_bounds__(kThreads) void k_fasta_write(const u8 *__restrict__ raw, u64 n,
                                                          u64 n_blocks,
                                                          const u8 *__restrict__ states,
                                                          const u64 *__restrict__ incl,
                                                          char *__restrict__ chars,
                                                          u64 *__restrict__ ends) {
  __shared__ u32 sh[kThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(raw, n, blk * kThreads + threadIdx.x, b);
    const u32 st = automaton::segment_start(blk, states, sh, b, len, FastaMap{});
    u32 nc = 0, nh = 0, plus = 0;
    walk<false>(b, len, st, nc, nh, plus, nullptr, 0, nullptr, 0);
    u32 tc, th;
    const u32 c_before = block_prefix_sum(nc, sh, &tc);
    const u32 h_before = block_prefix_sum(nh, sh, &th);
    const u64 base = blk ? incl[blk - 1] : 0;
    nc = 0;
    nh = 0;
    walk<true>(b, len, st, nc, nh, plus, chars, (base >> 32) + c_before, ends,
               (base & 0xFFFFFFFFu) + h_before);
  }
}

// The last read ends at the chunk's last char.
__global__ void k_fasta_last_end(u64 *ends, u64 n_reads, u64 n_chars) { ends[n_reads - 1] = n_chars; }

__global__ void k_fill_ones(u8 *p, u64 n) {
  for (u64 i = blockIdx.x * u64{blockDim.x} + threadIdx.x; i < n; i += u64{gridDim.x} * blockDim.x)
    p[i] = 1;
}

u32 grid_of(u64 n_blocks) { return static_cast<u32>(std::min<u64>(n_blocks, 1u << 16)); }

// ---- FASTQ: the state is the line's index within its record (mod 4)
constexpr u32 kMapNextLine = 1u | 2u << 2 | 3u << 4 | 0u << 6;
constexpr u32 kSeqLine = 1;

struct FastqMap {
  __device__ __forceinline__ u32 operator()(u8 b) const {
    return b == '\n' ? kMapNextLine : automaton::kMapIdentity;
  }
};

__global__ __launch_bounds__(kThreads) void k_fastq_maps(const u8 *__restrict__ raw, u64 n,
                                                         u64 n_blocks, u8 *__restrict__ maps) {
  automaton::block_maps(raw, n, n_blocks, maps, FastqMap{});
}

__global__ __launch_bounds__(kStatesThreads) void k_fastq_states(const u8 *__restrict__ maps,
                                                                 u64 n_blocks,
                                                                 u8 *__restrict__ states) {
  automaton::block_states(maps, n_blocks, 0, states);
}

// Per block: sequence chars << 32 | newlines.
__global__ __launch_bounds__(kThreads) void k_fastq_count(const u8 *__restrict__ raw, u64 n,
                                                          u64 n_blocks,
                                                          const u8 *__restrict__ states,
                                                          u64 *__restrict__ counts) {
  __shared__ u32 sh[kThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(raw, n, blk * kThreads + threadIdx.x, b);
    u32 st = automaton::segment_start(blk, states, sh, b, len, FastqMap{});
    u32 nc = 0, nl = 0;
#pragma unroll
    for (u32 i = 0; i < kSegBytes; ++i) {
      if (i < len) {
        const bool newline = b[i] == '\n';
        nc += (st == kSeqLine && !newline) ? 1u : 0u;
        nl += newline ? 1u : 0u;
        st = newline ? (st + 1) & 3u : st;
      }
    }
    u32 tc, tl;
    block_prefix_sum(nc, sh, &tc);
    block_prefix_sum(nl, sh, &tl);
    if (threadIdx.x == 0) counts[blk] = (u64{tc} << 32) | tl;
  }
}

// Copies the sequence chars; writes each line's '\n' offset (lines below
// line_cap) and each record's read end (records below line_cap / 4).
__global__ __launch_bounds__(kThreads) void k_fastq_write(const u8 *__restrict__ raw, u64 n,
                                                          u64 n_blocks,
                                                          const u8 *__restrict__ states,
                                                          const u64 *__restrict__ incl,
                                                          u64 line_cap, char *__restrict__ chars,
                                                          u64 *__restrict__ ends,
                                                          u32 *__restrict__ line_ends) {
  __shared__ u32 sh[kThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const u64 seg = blk * kThreads + threadIdx.x;
    u8 b[kSegBytes];
    const u32 len = load_segment(raw, n, seg, b);
    const u32 st0 = automaton::segment_start(blk, states, sh, b, len, FastqMap{});
    u32 st = st0, nc = 0, nl = 0;
#pragma unroll
    for (u32 i = 0; i < kSegBytes; ++i) {
      if (i < len) {
        const bool newline = b[i] == '\n';
        nc += (st == kSeqLine && !newline) ? 1u : 0u;
        nl += newline ? 1u : 0u;
        st = newline ? (st + 1) & 3u : st;
      }
    }
    u32 tc, tl;
    const u32 c_before = block_prefix_sum(nc, sh, &tc);
    const u32 l_before = block_prefix_sum(nl, sh, &tl);
    const u64 base = blk ? incl[blk - 1] : 0;
    const u64 cbase = (base >> 32) + c_before;
    const u64 lbase = (base & 0xFFFFFFFFu) + l_before;
    st = st0;
    nc = 0;
    nl = 0;
#pragma unroll
    for (u32 i = 0; i < kSegBytes; ++i) {
      if (i < len) {
        const bool newline = b[i] == '\n';
        if (st == kSeqLine && !newline) chars[cbase + nc] = static_cast<char>(b[i]);
        if (newline) {
          const u64 line = lbase + nl;
          if (line < line_cap) {
            line_ends[line] = static_cast<u32>(seg * kSegBytes + i);
            if (st == kSeqLine) ends[line >> 2] = cbase + nc;
          }
        }
        nc += (st == kSeqLine && !newline) ? 1u : 0u;
        nl += newline ? 1u : 0u;
        st = newline ? (st + 1) & 3u : st;
      }
    }
  }
}

__global__ void k_set_u32(u32 *p, u32 v) { *p = v; }

__device__ __forceinline__ bool header_start(u8 c) { return c == '@' || c == '>'; }

// Record r (below n_rec) reads as the record reader reads it: see seq_parse.hpp.
__global__ void k_fastq_check(const u8 *__restrict__ raw, const u32 *__restrict__ line_ends,
                              u64 n_rec, u32 *first_bad) {
  for (u64 r = blockIdx.x * u64{blockDim.x} + threadIdx.x; r < n_rec;
       r += u64{gridDim.x} * blockDim.x) {
    const u64 s0 = r ? u64{line_ends[4 * r - 1]} + 1 : 0;
    const u64 s1 = u64{line_ends[4 * r]} + 1;
    const u64 s2 = u64{line_ends[4 * r + 1]} + 1;
    const u64 s3 = u64{line_ends[4 * r + 2]} + 1;
    const u64 len1 = s2 - 1 - s1, len3 = u64{line_ends[4 * r + 3]} - s3;
    const u8 c1 = raw[s1];
    const bool ok = header_start(raw[s0]) && len1 > 0 && c1 != '+' && !header_start(c1) &&
                    raw[s2] == '+' && len3 >= len1;
    if (!ok) atomicMin(first_bad, static_cast<u32>(r));
  }
}

// out: records parsed, prefix bytes, chars.
__global__ void k_fastq_finish(const u32 *first_bad, const u32 *line_ends, const u64 *ends, u64 n,
                               u64 *out) {
  const u64 r = *first_bad;
  out[0] = r;
  out[1] = r ? min<u64>(u64{line_ends[4 * r - 1]} + 1, n) : 0;
  out[2] = r ? ends[r - 1] : 0;
}

}  // namespace

FastaParser::FastaParser(u64 max_bytes, u64 max_reads, hipStream_t stream)
    : max_bytes_(max_bytes), max_reads_(max_reads), stream_(stream), own_stream_(stream == nullptr) {
  if (max_bytes_ == 0 || max_bytes_ >= (1ULL << 32))
    throw std::runtime_error("FASTA chunks must hold between 1 byte and 4 GiB - 1");
  if (max_reads_ == 0) throw std::runtime_error("max_reads must be positive");
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  block_cap_ = max_bytes_ / kBlockBytes + 2;
  SBWT_HIP(hipMalloc(&d_raw_, max_bytes_ + 64));
  SBWT_HIP(hipMalloc(&d_chars_, max_bytes_ + 64));
  SBWT_HIP(hipMalloc(&d_ends_, max_reads_ * 8));
  SBWT_HIP(hipMalloc(&d_ones_, max_reads_));
  SBWT_HIP(hipMalloc(&d_block_maps_, block_cap_ * 2));
  SBWT_HIP(hipMalloc(&d_block_counts_, block_cap_ * 2 * 8));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(block_cap_) * 8));
  SBWT_HIP(hipMalloc(&d_flags_, 4));
  k_fill_ones<<<static_cast<u32>(std::min<u64>((max_reads_ + 255) / 256, 1u << 16)), 256, 0, stream_>>>(
      d_ones_, max_reads_);
  SBWT_HIP(hipGetLastError());
}

FastaParser::~FastaParser() {
  void *bufs[] = {d_raw_, d_chars_, d_ends_, d_ones_, d_block_maps_, d_block_counts_, d_scan_tmp_,
                  d_flags_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

u64 FastaParser::parse_host(const char *raw, u64 n_bytes, u64 *n_chars) {
  *n_chars = 0;
  if (n_bytes == 0) return 0;
  if (n_bytes > max_bytes_) throw std::runtime_error("FASTA chunk exceeds the parser's max bytes");
  if (raw[0] != '>' && raw[0] != '@') throw std::runtime_error("FASTA chunk must start with a header line");
  SBWT_HIP(hipMemcpyAsync(d_raw_, raw, n_bytes, hipMemcpyHostToDevice, stream_));
  const u8 *body = reinterpret_cast<const u8 *>(d_raw_);
  const u64 n_blocks = (n_bytes + kBlockBytes - 1) / kBlockBytes;
  u8 *states = d_block_maps_ + block_cap_;
  u64 *incl = d_block_counts_ + block_cap_;
  const u32 grid = grid_of(n_blocks);
  SBWT_HIP(hipMemsetAsync(d_flags_, 0, 4, stream_));
  k_fasta_maps<<<grid, kThreads, 0, stream_>>>(body, n_bytes, n_blocks, d_block_maps_);
  k_fasta_states<<<1, kStatesThreads, 0, stream_>>>(d_block_maps_, n_blocks, states);
  k_fasta_count<<<grid, kThreads, 0, stream_>>>(body, n_bytes, n_blocks, states, d_block_counts_,
                                                d_flags_);
  SBWT_HIP(hipGetLastError());
  scan_inclusive_u64(d_block_counts_, n_blocks, incl, d_scan_tmp_, stream_);
  u64 tot = 0;
  u32 flags = 0;
  SBWT_HIP(hipMemcpyAsync(&tot, incl + n_blocks - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipMemcpyAsync(&flags, d_flags_, 4, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  if (flags) throw std::runtime_error("FASTA chunk holds a line starting with '+' (FASTQ quality)");
  const u64 nc = tot >> 32, nh = tot & 0xFFFFFFFFu;
  if (nh > max_reads_) throw std::runtime_error("FASTA chunk holds more records than the parser's max reads");
  k_fasta_write<<<grid, kThreads, 0, stream_>>>(body, n_bytes, n_blocks, states, incl, d_chars_,
                                                d_ends_);
  k_fasta_last_end<<<1, 1, 0, stream_>>>(d_ends_, nh, nc);
  SBWT_HIP(hipGetLastError());
  *n_chars = nc;
  return nh;
}

FastqParser::FastqParser(u64 max_bytes, u64 max_reads, hipStream_t stream)
    : max_bytes_(max_bytes), max_reads_(max_reads), stream_(stream), own_stream_(stream == nullptr) {
  if (max_bytes_ == 0 || max_bytes_ >= (1ULL << 32))
    throw std::runtime_error("FASTQ chunks must hold between 1 byte and 4 GiB - 1");
  if (max_reads_ == 0 || max_reads_ >= (1ULL << 30)) throw std::runtime_error("max_reads must be in [1, 2^30)");
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  block_cap_ = max_bytes_ / kBlockBytes + 2;
  SBWT_HIP(hipMalloc(&d_raw_, max_bytes_ + 64));
  SBWT_HIP(hipMalloc(&d_chars_, max_bytes_ + 64));
  SBWT_HIP(hipMalloc(&d_ends_, max_reads_ * 8));
  SBWT_HIP(hipMalloc(&d_ones_, max_reads_));
  SBWT_HIP(hipMalloc(&d_line_ends_, max_reads_ * 4 * 4));
  SBWT_HIP(hipMalloc(&d_block_maps_, block_cap_ * 2));
  SBWT_HIP(hipMalloc(&d_block_counts_, block_cap_ * 2 * 8));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(block_cap_) * 8));
  SBWT_HIP(hipMalloc(&d_first_bad_, 4));
  SBWT_HIP(hipMalloc(&d_result_, 3 * 8));
  k_fill_ones<<<static_cast<u32>(std::min<u64>((max_reads_ + 255) / 256, 1u << 16)), 256, 0, stream_>>>(
      d_ones_, max_reads_);
  SBWT_HIP(hipGetLastError());
}

FastqParser::~FastqParser() {
  void *bufs[] = {d_raw_, d_chars_, d_ends_, d_ones_, d_line_ends_, d_block_maps_, d_block_counts_,
                  d_scan_tmp_, d_first_bad_, d_result_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

u64 FastqParser::parse_host(const char *raw, u64 n, bool last, u64 *n_reads, u64 *n_chars) {
  *n_reads = 0;
  *n_chars = 0;
  if (n == 0) return 0;
  if (n > max_bytes_) throw std::runtime_error("FASTQ chunk exceeds the parser's max bytes");
  SBWT_HIP(hipMemcpyAsync(d_raw_, raw, n, hipMemcpyHostToDevice, stream_));
  const u8 *body = reinterpret_cast<const u8 *>(d_raw_);
  const u64 n_blocks = (n + kBlockBytes - 1) / kBlockBytes;
  u8 *states = d_block_maps_ + block_cap_;
  u64 *incl = d_block_counts_ + block_cap_;
  const u32 grid = grid_of(n_blocks);
  k_fastq_maps<<<grid, kThreads, 0, stream_>>>(body, n, n_blocks, d_block_maps_);
  k_fastq_states<<<1, kStatesThreads, 0, stream_>>>(d_block_maps_, n_blocks, states);
  k_fastq_count<<<grid, kThreads, 0, stream_>>>(body, n, n_blocks, states, d_block_counts_);
  SBWT_HIP(hipGetLastError());
  scan_inclusive_u64(d_block_counts_, n_blocks, incl, d_scan_tmp_, stream_);
  u64 tot = 0;
  SBWT_HIP(hipMemcpyAsync(&tot, incl + n_blocks - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  const u64 newlines = tot & 0xFFFFFFFFu;
  const bool open_line = last && raw[n - 1] != '\n';  // a last line without '\n'
  const u64 lines = newlines + (open_line ? 1 : 0);
  const u64 r_cap = std::min<u64>(lines / 4, max_reads_);
  if (r_cap == 0) return 0;
  const u64 line_cap = 4 * r_cap;
  k_fastq_write<<<grid, kThreads, 0, stream_>>>(body, n, n_blocks, states, incl, line_cap,
                                                d_chars_, d_ends_, d_line_ends_);
  if (open_line && newlines < line_cap)
    k_set_u32<<<1, 1, 0, stream_>>>(d_line_ends_ + newlines, static_cast<u32>(n));
  k_set_u32<<<1, 1, 0, stream_>>>(d_first_bad_, static_cast<u32>(r_cap));
  k_fastq_check<<<static_cast<u32>(std::min<u64>((r_cap + 255) / 256, 1u << 16)), 256, 0, stream_>>>(
      body, d_line_ends_, r_cap, d_first_bad_);
  k_fastq_finish<<<1, 1, 0, stream_>>>(d_first_bad_, d_line_ends_, d_ends_, n, d_result_);
  SBWT_HIP(hipGetLastError());
  u64 res[3] = {0, 0, 0};
  SBWT_HIP(hipMemcpyAsync(res, d_result_, sizeof(res), hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  *n_reads = res[0];
  *n_chars = res[2];
  return res[1];
}

u64 FastaParser::chunk_cut(const char *raw, u64 n, bool last, u64 max_reads) {
  if (n == 0 || max_reads == 0) return 0;
  // Header starts are position 0 and every '>' or '@' right after a '\n'; two
  // of them are at least 2 bytes apart, so raw holds at most (n + 1) / 2.
  u64 limit = n;  // the start of record max_reads + 1, if raw holds one
  if ((n + 1) / 2 > max_reads) {
    u64 seen = 1;
    const char *p = raw, *end = raw + n;
    while ((p = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)))) &&
           ++p < end) {
      if ((*p == '>' || *p == '@') && ++seen > max_reads) {
        limit = static_cast<u64>(p - raw);
        break;
      }
    }
  }
  if (last && limit == n) return n;
  // the last header start in [1, min(limit, n - 1)]: a cut at n itself needs
  // the next record's header to follow, which only `last` tells
  u64 len = std::min(limit, n - 1);  // '\n' searched in raw[0, len)
  while (len > 0) {
    const char *nl = static_cast<const char *>(memrchr(raw, '\n', static_cast<size_t>(len)));
    if (!nl) break;
    const u64 at = static_cast<u64>(nl - raw) + 1;
    if (raw[at] == '>' || raw[at] == '@') return at;
    len = static_cast<u64>(nl - raw);
  }
  return 0;
}

bool FastaParser::has_quality_line(const char *raw, u64 n, bool at_line_start) {
  if (n == 0) return false;
  if (at_line_start && raw[0] == '+') return true;
  const char *p = raw, *end = raw + n;
  while ((p = static_cast<const char *>(std::memchr(p, '+', static_cast<size_t>(end - p))))) {
    if (p > raw && p[-1] == '\n') return true;
    ++p;
  }
  return false;
}

}  // namespace sbwt_amd
