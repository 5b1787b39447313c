// Device parser of index-result files (see index_parse.hpp).
// Ascii: four kernels over 4 KiB blocks of 256 threads x 16-byte segments:
//   maps    each block's state map (composition of its bytes' maps)
//   states  one workgroup scans the block maps into each block's start state
//   count   per block: results and read ends (each thread walks its segment
//           from its start state: the block's start state through the
//           exclusive prefix of its segment maps)
//   write   the same walk, writing at offsets from the scan of the counts
// Binary: count and write kernels over u64 words.
#include "index_parse.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "byte_automaton.hpp"
#include "host_scan.hpp"
#include "scan.hpp"

namespace sbwt_amd {
namespace {

using automaton::block_prefix_sum;
using automaton::kBlockBytes;
using automaton::kSegBytes;
using automaton::kStatesThreads;
using automaton::load_segment;
using automaton::next_state;
constexpr u32 kParseThreads = automaton::kThreads;
constexpr u64 kBinaryBlockBytes = u64{kParseThreads} * 8;  // binary block: 256 words

// Automaton states: top level, inside a number, after '-', skipping the
// digits after the char that follows '-'.
constexpr u32 kTop = 0, kNum = 1, kDash = 2, kSkip = 3;
// A state map packs the next state of start state s in bits [2s, 2s + 2).
constexpr u32 kMapDigit = kNum | kNum << 2 | kSkip << 4 | kSkip << 6;
constexpr u32 kMapMinus = kDash | kDash << 2 | kSkip << 4 | kTop << 6;
constexpr u32 kMapOther = kTop | kTop << 2 | kSkip << 4 | kTop << 6;  // '\n' and the rest

__device__ __forceinline__ bool is_digit(u8 b) { return static_cast<signed char>(b) >= '0'; }

struct AsciiMap {
  __device__ __forceinline__ u32 operator()(u8 b) const {
    return is_digit(b) ? kMapDigit : (b == '-' ? kMapMinus : kMapOther);
  }
};

// u64 value of the digit run at body[pos] (reference parse_number: wrapping
// arithmetic over every char >= '0').
__device__ __forceinline__ u64 parse_number(const u8 *body, u64 n, u64 pos) {
  u64 v = 0;
  for (u64 j = pos; j < n; ++j) {
    const u8 c = body[j];
    if (!is_digit(c)) break;
    v = v * 10 + static_cast<u64>(static_cast<signed char>(c) - '0');
  }
  return v;
}

// Runs the automaton over one segment from state st, counting results (r)
// and read ends (nl); with WRITE, also stores them at rbase + r and
// nlbase + nl.
template <bool WRITE>
__device__ __forceinline__ void walk_segment(const u8 *body, u64 n, u64 beg, const u8 (&b)[kSegBytes],
                                             u32 len, u32 st, u32 &r, u32 &nl, u64 *results,
                                             u64 rbase, u64 *kends, u64 nlbase) {
#pragma unroll
  for (u32 i = 0; i < kSegBytes; ++i) {
    if (i < len) {
      const u8 c = b[i];
      // what this char emits: a result (counted '-1' / '-2', or a number's
      // first digit) or a read end
      const bool dash_result = st == kDash && (c == '1' || c == '2');
      const bool number = st == kTop && is_digit(c);
      const bool read_end = st != kDash && c == '\n';
      if constexpr (WRITE) {
        if (dash_result) results[rbase + r] = c == '1' ? kNotFound : kInvalid;
        if (number) results[rbase + r] = parse_number(body, n, beg + i);
        if (read_end) kends[nlbase + nl] = rbase + r;
      }
      r += (dash_result || number) ? 1u : 0u;
      nl += read_end ? 1u : 0u;
      st = next_state(AsciiMap{}(c), st);
    }
  }
}

__global__ __launch_bounds__(kParseThreads) void k_ascii_maps(const u8 *__restrict__ body, u64 n,
                                                              u64 n_blocks, u8 *__restrict__ maps) {
  automaton::block_maps(body, n, n_blocks, maps, AsciiMap{});
}

__global__ __launch_bounds__(kStatesThreads) void k_ascii_states(const u8 *__restrict__ maps,
                                                                 u64 n_blocks,
                                                                 u8 *__restrict__ states) {
  automaton::block_states(maps, n_blocks, kTop, states);
}

__global__ __launch_bounds__(kParseThreads) void k_ascii_count(const u8 *__restrict__ body, u64 n,
                                                               u64 n_blocks,
                                                               const u8 *__restrict__ states,
                                                               u64 *__restrict__ counts) {
  __shared__ u32 sh[kParseThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(body, n, blk * kParseThreads + threadIdx.x, b);
    const u32 st = automaton::segment_start(blk, states, sh, b, len, AsciiMap{});
    u32 r = 0, nl = 0;
    const u64 beg = (blk * kParseThreads + threadIdx.x) * kSegBytes;
    walk_segment<false>(body, n, beg, b, len, st, r, nl, nullptr, 0, nullptr, 0);
    u32 tr, tn;
    block_prefix_sum(r, sh, &tr);
    block_prefix_sum(nl, sh, &tn);
    if (threadIdx.x == 0) counts[blk] = (u64{tr} << 32) | tn;
  }
}

__global__ __launch_bounds__(kParseThreads) void k_ascii_write(const u8 *__restrict__ body, u64 n,
                                                               u64 n_blocks,
                                                               const u8 *__restrict__ states,
                                                               const u64 *__restrict__ incl,
                                                               u64 *__restrict__ results,
                                                               u64 *__restrict__ kends) {
  __shared__ u32 sh[kParseThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    u8 b[kSegBytes];
    const u32 len = load_segment(body, n, blk * kParseThreads + threadIdx.x, b);
    const u32 st = automaton::segment_start(blk, states, sh, b, len, AsciiMap{});
    u32 r = 0, nl = 0;
    const u64 beg = (blk * kParseThreads + threadIdx.x) * kSegBytes;
    walk_segment<false>(body, n, beg, b, len, st, r, nl, nullptr, 0, nullptr, 0);
    u32 tr, tn;
    const u32 r_before = block_prefix_sum(r, sh, &tr);
    const u32 nl_before = block_prefix_sum(nl, sh, &tn);
    const u64 base = blk ? incl[blk - 1] : 0;
    r = 0;
    nl = 0;
    walk_segment<true>(body, n, beg, b, len, st, r, nl, results, (base >> 32) + r_before, kends,
                       (base & 0xFFFFFFFFu) + nl_before);
  }
}

__global__ __launch_bounds__(kParseThreads) void k_binary_count(const u64 *__restrict__ words,
                                                                u64 n_words, u64 n_blocks,
                                                                u64 *__restrict__ counts) {
  __shared__ u32 sh[kParseThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const u64 i = blk * kParseThreads + threadIdx.x;
    const u32 nl = i < n_words && words[i] == kNewline ? 1u : 0u;
    const u32 r = i < n_words && !nl ? 1u : 0u;
    u32 tr, tn;
    block_prefix_sum(r, sh, &tr);
    block_prefix_sum(nl, sh, &tn);
    if (threadIdx.x == 0) counts[blk] = (u64{tr} << 32) | tn;
  }
}

__global__ __launch_bounds__(kParseThreads) void k_binary_write(const u64 *__restrict__ words,
                                                                u64 n_words, u64 n_blocks,
                                                                const u64 *__restrict__ incl,
                                                                u64 *__restrict__ results,
                                                                u64 *__restrict__ kends) {
  __shared__ u32 sh[kParseThreads / 64];
  for (u64 blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const u64 i = blk * kParseThreads + threadIdx.x;
    const u64 w = i < n_words ? words[i] : 0;
    const u32 nl = i < n_words && w == kNewline ? 1u : 0u;
    const u32 r = i < n_words && !nl ? 1u : 0u;
    u32 tr, tn;
    const u32 r_before = block_prefix_sum(r, sh, &tr);
    const u32 nl_before = block_prefix_sum(nl, sh, &tn);
    const u64 base = blk ? incl[blk - 1] : 0;
    const u64 ri = (base >> 32) + r_before;
    if (r) results[ri] = w;
    if (nl) kends[(base & 0xFFFFFFFFu) + nl_before] = ri;
  }
}

u32 grid_of(u64 n_blocks) { return static_cast<u32>(std::min<u64>(n_blocks, 1u << 16)); }

}  // namespace

IndexFileParser::IndexFileParser(u64 max_bytes, u64 max_reads, hipStream_t stream)
    : max_bytes_(max_bytes), max_reads_(max_reads), stream_(stream), own_stream_(stream == nullptr) {
  if (max_bytes_ == 0 || max_bytes_ >= (1ULL << 32))
    throw std::runtime_error("index file batches must hold between 1 byte and 4 GiB - 1");
  // every ascii result starts at its own byte and the next one starts at
  // least two bytes later; a binary result is 8 bytes
  max_results_ = max_bytes_ / 2 + 1;
  if (own_stream_) SBWT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // binary blocks are the smaller, so they bound the block count
  block_cap_ = max_bytes_ / kBinaryBlockBytes + 2;
  const u64 blocks = block_cap_;
  SBWT_HIP(hipMalloc(&d_body_, max_bytes_ + 64));
  SBWT_HIP(hipMalloc(&d_results_, max_results_ * 8));
  SBWT_HIP(hipMalloc(&d_kends_, std::max<u64>(max_reads_, 1) * 8));
  SBWT_HIP(hipMalloc(&d_block_maps_, blocks * 2));
  SBWT_HIP(hipMalloc(&d_block_counts_, blocks * 2 * 8));
  SBWT_HIP(hipMalloc(&d_scan_tmp_, scan_tmp_entries(blocks) * 8));
}

IndexFileParser::~IndexFileParser() {
  void *bufs[] = {d_body_, d_results_, d_kends_, d_block_maps_, d_block_counts_, d_scan_tmp_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

u64 IndexFileParser::parse_device(int format, const char *d_body, u64 n_bytes, u64 *n_results) {
  *n_results = 0;
  if (format != kIndexFileAscii && format != kIndexFileBinary)
    throw std::runtime_error("index file format must be ascii or binary");
  if (n_bytes > max_bytes_) throw std::runtime_error("index file batch exceeds the parser's max bytes");
  if (reinterpret_cast<uintptr_t>(d_body) % 16)
    throw std::runtime_error("index file body must be 16-byte aligned on the device");
  const u8 *body = reinterpret_cast<const u8 *>(d_body);
  const u64 units = format == kIndexFileAscii ? n_bytes : n_bytes / 8;
  if (units == 0) return 0;
  const u64 n_blocks = format == kIndexFileAscii ? (n_bytes + kBlockBytes - 1) / kBlockBytes
                                                 : (units + kParseThreads - 1) / kParseThreads;
  if (n_blocks > block_cap_) throw std::runtime_error("index file batch exceeds the parser's blocks");
  u8 *states = d_block_maps_ + block_cap_;
  u64 *incl = d_block_counts_ + block_cap_;
  const u32 grid = grid_of(n_blocks);
  if (format == kIndexFileAscii) {
    k_ascii_maps<<<grid, kParseThreads, 0, stream_>>>(body, n_bytes, n_blocks, d_block_maps_);
    k_ascii_states<<<1, kStatesThreads, 0, stream_>>>(d_block_maps_, n_blocks, states);
    k_ascii_count<<<grid, kParseThreads, 0, stream_>>>(body, n_bytes, n_blocks, states,
                                                       d_block_counts_);
  } else {
    k_binary_count<<<grid, kParseThreads, 0, stream_>>>(reinterpret_cast<const u64 *>(body), units,
                                                        n_blocks, d_block_counts_);
  }
  SBWT_HIP(hipGetLastError());
  scan_inclusive_u64(d_block_counts_, n_blocks, incl, d_scan_tmp_, stream_);
  u64 tot = 0;
  SBWT_HIP(hipMemcpyAsync(&tot, incl + n_blocks - 1, 8, hipMemcpyDeviceToHost, stream_));
  SBWT_HIP(hipStreamSynchronize(stream_));
  const u64 nr = tot >> 32, nl = tot & 0xFFFFFFFFu;
  if (nl > max_reads_) throw std::runtime_error("index file batch holds more reads than the parser's max reads");
  if (nr > max_results_) throw std::runtime_error("index file batch holds more results than the parser can");
  if (format == kIndexFileAscii) {
    k_ascii_write<<<grid, kParseThreads, 0, stream_>>>(body, n_bytes, n_blocks, states, incl,
                                                       d_results_, d_kends_);
  } else {
    k_binary_write<<<grid, kParseThreads, 0, stream_>>>(reinterpret_cast<const u64 *>(body), units,
                                                        n_blocks, incl, d_results_, d_kends_);
  }
  SBWT_HIP(hipGetLastError());
  *n_results = nr;
  return nl;
}

u64 IndexFileParser::parse_host(int format, const char *body, u64 n_bytes, u64 *n_results) {
  *n_results = 0;
  if (n_bytes > max_bytes_) throw std::runtime_error("index file batch exceeds the parser's max bytes");
  if (n_bytes)
    SBWT_HIP(hipMemcpyAsync(d_body_, body, n_bytes, hipMemcpyHostToDevice, stream_));
  return parse_device(format, d_body_, n_bytes, n_results);
}

namespace {

// Start of the scan for the max_reads-th read end among n units (bytes or
// words): a batch of at least kParallelUnits units is counted first in 64
// parts by four threads (count(b, e) = read ends in units [b, e)). Returns
// false when the batch holds fewer than max_reads read ends, so that the cap
// cannot bind; otherwise *start is the first unit of the part holding the
// max_reads-th one and *seen the read ends before it (both 0 for a small
// batch).
constexpr u64 kParallelUnits = 1ULL << 20;

template <class Count>
bool cap_scan_start(u64 n, u64 max_reads, Count count, u64 *start, u64 *seen) {
  *start = 0;
  *seen = 0;
  if (n <= max_reads) return false;
  if (n < kParallelUnits) return true;
  constexpr u32 kThreads = 4, kParts = 64, kPartsPerThread = kParts / kThreads;
  u64 part_counts[kParts] = {};
  std::thread th[kThreads];
  const u64 step = (n + kParts - 1) / kParts;
  for (u32 t = 0; t < kThreads; ++t)
    th[t] = std::thread([&, t] {
      for (u32 q = t * kPartsPerThread; q < (t + 1) * kPartsPerThread; ++q) {
        const u64 b = std::min(n, q * step);
        part_counts[q] = count(b, std::min(n, b + step));
      }
    });
  for (u32 t = 0; t < kThreads; ++t) th[t].join();
  for (u32 t = 0; t < kParts; ++t) {
    if (*seen + part_counts[t] >= max_reads) {
      *start = std::min(n, t * step);
      return true;
    }
    *seen += part_counts[t];
  }
  return false;
}

}  // namespace

u64 IndexFileParser::colored_cut(int format, const char *body, u64 n) {
  if (format == kIndexFileBinary) {
    // a word below u64 max - 2 is a found index
    u64 end = 0;  // words of the candidate read end (exclusive), 0: none yet
    bool found = false;
    for (u64 w = n / 8; w > 0; --w) {
      u64 v;
      std::memcpy(&v, body + 8 * (w - 1), 8);
      if (v == kNewline) {
        if (end && found) return 8 * end;
        end = w;
        found = false;
      } else if (end && v < kNewline) {
        found = true;
      }
    }
    return end && found ? 8 * end : 0;
  }
  // ascii: read ends are '\n' not preceded by '-' (as batch_cut); a read
  // has a found index when the reference's loop meets a char >= '0' outside
  // the skip that follows a '-'
  const auto line_found = [&](u64 s, u64 e) {
    for (u64 i = s; i < e;) {
      const signed char c = static_cast<signed char>(body[i++]);
      if (c == '-') {
        ++i;  // the char after '-' (1: not found, 2: invalid, else nothing)
        while (i < e && static_cast<signed char>(body[i]) >= '0') ++i;
        ++i;  // the char that ended the skip
        continue;
      }
      if (c >= '0') return true;
    }
    return false;
  };
  u64 c = n;
  u64 end = 0;  // candidate: offset right after a read end
  while (c > 0) {
    const void *nl = memrchr(body, '\n', static_cast<size_t>(c));
    u64 at = nl ? static_cast<u64>(static_cast<const char *>(nl) - body) : 0;
    if (nl && at > 0 && body[at - 1] == '-') {  // consumed after a '-': not a read end
      c = at;
      continue;
    }
    const u64 line_start = nl ? at + 1 : 0;
    if (end && line_found(line_start, end - 1)) return end;
    if (!nl) break;
    end = at + 1;
    c = at;
  }
  return 0;
}

u64 IndexFileParser::batch_cut(int format, const char *body, u64 n, u64 max_reads) {
  if (max_reads == 0) return 0;
  if (format == kIndexFileBinary) {
    // read ends are u64 max - 2 words: stop after the max_reads-th one, then
    // step back to the last one
    const u64 words = n / 8;
    u64 limit = words, start, seen;
    const auto count = [&](u64 b, u64 e) { return count_word(body + 8 * b, e - b, kNewline); };
    if (cap_scan_start(words, max_reads, count, &start, &seen)) {
      for (u64 w = start; w < words;) {
        const u64 span = std::min<u64>(words - w, 1 << 16);
        const u64 c = count_word(body + 8 * w, span, kNewline);
        if (seen + c < max_reads) {
          seen += c;
          w += span;
          continue;
        }
        for (u64 q = w; q < w + span; ++q) {
          u64 v;
          std::memcpy(&v, body + 8 * q, 8);
          if (v == kNewline && ++seen == max_reads) {
            limit = q + 1;
            break;
          }
        }
        break;
      }
    }
    for (u64 w = limit; w > 0; --w) {
      u64 v;
      std::memcpy(&v, body + 8 * (w - 1), 8);
      if (v == kNewline) return 8 * w;
    }
    return 0;
  }
  // A read end is at most one per '\n' byte: stop at the max_reads-th one,
  // then step back to a '\n' that is not the char consumed after a '-'.
  u64 limit = n;
  u64 start, seen;
  const auto count = [&](u64 b, u64 e) { return count_byte(body + b, e - b, '\n'); };
  if (cap_scan_start(n, max_reads, count, &start, &seen)) {
    for (u64 p = start; p < n;) {
      const u64 span = std::min<u64>(n - p, 1 << 16);
      const u64 c = count_byte(body + p, span, '\n');
      if (seen + c < max_reads) {
        seen += c;
        p += span;
        continue;
      }
      for (u64 q = p; q < p + span; ++q)
        if (body[q] == '\n' && ++seen == max_reads) {
          limit = q + 1;
          break;
        }
      break;
    }
  }
  for (u64 c = limit; c > 0;) {
    const void *nl = memrchr(body, '\n', static_cast<size_t>(c));
    if (!nl) break;
    c = static_cast<u64>(static_cast<const char *>(nl) - body) + 1;
    if (c < 2 || body[c - 2] != '-') return c;
    --c;
  }
  return 0;
}

}  // namespace sbwt_amd
