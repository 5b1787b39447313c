"""ctypes binding of libsbwt_amd.so (the C ABI in include/sbwt_amd.h).

This is the Python-side FFI stub of the engine: it loads the in-tree shared
library and raises if it is missing -- there is no CPU fallback. Device
pointers are plain integers (e.g. torch.Tensor.data_ptr()).
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "csrc", "libsbwt_amd.so")
HEADER = os.path.join(REPO, "include", "sbwt_amd.h")

ALGO_STREAMING = 0
ALGO_PER_KMER = 1
ALGO_AUTO = 2  # default: per k-mer below 2^22 chars per batch, streaming (adaptive work items) above
ALGO_STREAMING_ADAPTIVE = 3  # adaptive work items from 2^16 chars: batches sharing the GPU
COLOR_OUTPUT_NONE, COLOR_OUTPUT_ENTRIES, COLOR_OUTPUT_CHAINS = 0, 1, 2  # sbwt_searcher_set_color_output
NOALLOC_MAX_CHARS = 1 << 22
NOALLOC_MAX_READS = 1 << 18
NOT_FOUND = 0xFFFFFFFFFFFFFFFF
INVALID = 0xFFFFFFFFFFFFFFFE
WALK_STUCK = 0xFFFFFFFFFFFFFFFC

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class IndexInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint64), ("kmer_size", C.c_uint32),
                ("presearch_letters", C.c_uint32), ("c_map", C.c_uint64 * 5),
                ("device_bytes", C.c_uint64), ("has_key_kmer_marks", C.c_int), ("wide", C.c_int),
                ("has_suffix_groups", C.c_int), ("suffix_groups_rejected", C.c_int)]


class ColorIndexInfo(C.Structure):
    _fields_ = [("num_colors", C.c_uint64), ("num_color_sets", C.c_uint64),
                ("num_nodes", C.c_uint64), ("device_bytes", C.c_uint64)]


class ColorParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("include_not_found", C.c_int),
                ("include_invalid", C.c_int), ("first_read_quirk", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_chars_per_slot", C.c_uint64),
                ("max_reads_per_slot", C.c_uint64), ("memory_fraction", C.c_double),
                ("move_to_key_kmer", C.c_int), ("color", ColorParams), ("no_graphs", C.c_int),
                ("no_kmer_results", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("n_kmers", C.c_uint64), ("kmer_results", u64p), ("n_reads", C.c_uint64),
                ("color_offsets", u64p), ("color_ids", u32p), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("end_s", C.c_double)]


class SbwtError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libsbwt_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SbwtError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # One HIP runtime per process: torch ships its own libamdhip64 (same SONAME as
        # the system one the library links). Loading torch first makes the library
        # bind to that copy, so torch tensors and library kernels share one runtime.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sbwt_last_error.restype = C.c_char_p
        L.sbwt_mem_info.argtypes = [u64p, u64p]
        L.sbwt_random_line_probe.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_float), u64p, u64p]
        L.sbwt_index_create.argtypes = [u64p * 4, C.c_uint64, C.c_uint32, u64p, u64p,
                                        C.c_uint32, C.POINTER(vp)]
        L.sbwt_index_load.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(vp)]
        L.sbwt_index_build_kmer_table.argtypes = [vp, C.POINTER(C.c_int)]
        L.sbwt_suffix_groups_consistent.argtypes = [u64p * 4, u64p, C.c_uint64]
        L.sbwt_index_search_reads_noalloc.argtypes = [vp, C.c_int, vp, C.c_uint64, vp,
                                                      C.c_uint64, vp, vp]
        L.sbwt_index_destroy.argtypes = [vp]
        L.sbwt_index_destroy.restype = None
        L.sbwt_index_get_info.argtypes = [vp, C.POINTER(IndexInfo)]
        L.sbwt_searcher_create.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_int, vp, C.POINTER(vp)]
        L.sbwt_searcher_destroy.argtypes = [vp]
        L.sbwt_searcher_destroy.restype = None
        L.sbwt_searcher_sync.argtypes = [vp]
        L.sbwt_searcher_search.argtypes = [vp, u64p, C.c_uint64, u64p, C.c_uint64, u64p]
        L.sbwt_searcher_search_device.argtypes = [vp, vp, vp, C.c_uint64, vp]
        L.sbwt_searcher_search_reads.argtypes = [vp, C.c_char_p, C.c_uint64, u64p, C.c_uint64,
                                                 u64p, u64p]
        L.sbwt_searcher_search_reads_device.argtypes = [vp, vp, C.c_uint64, vp, C.c_uint64, vp,
                                                        u64p]
        L.sbwt_searcher_last_timings.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                 C.POINTER(C.c_float)]
        L.sbwt_pack_2bit_device.argtypes = [vp, C.c_uint64, vp, vp, vp]
        L.sbwt_searcher_kmer_ends_device.argtypes = [vp, C.POINTER(vp)]
        L.sbwt_formatter_create.argtypes = [C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_formatter_destroy.argtypes = [vp]
        L.sbwt_formatter_destroy.restype = None
        L.sbwt_formatter_format.argtypes = [vp, C.c_int, vp, vp, vp, C.c_uint64, C.c_int, vp,
                                            C.c_uint64, C.POINTER(C.c_uint64)]
        L.sbwt_color_index_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.sbwt_color_index_destroy.argtypes = [vp]
        L.sbwt_color_index_destroy.restype = None
        L.sbwt_color_index_get_info.argtypes = [vp, C.POINTER(ColorIndexInfo)]
        L.sbwt_color_searcher_create.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_color_searcher_destroy.argtypes = [vp]
        L.sbwt_color_searcher_destroy.restype = None
        L.sbwt_color_searcher_search.argtypes = [vp, u64p, C.c_uint64, u64p, C.c_uint64, u64p]
        L.sbwt_color_searcher_pseudoalign.argtypes = [vp, u64p, u64p, C.c_uint64,
                                                      C.POINTER(ColorParams), u64p, u32p,
                                                      C.c_uint64, u32p, u64p]
        L.sbwt_color_searcher_pseudoalign_device.argtypes = [vp, vp, vp, C.c_uint64,
                                                             C.POINTER(ColorParams), vp, vp,
                                                             C.c_uint64, vp, u64p]
        L.sbwt_color_searcher_counts_device.argtypes = [vp, vp, vp, C.c_uint64, vp]
        L.sbwt_color_searcher_set_work_items.argtypes = [vp, C.c_uint32]
        L.sbwt_color_searcher_autotune.argtypes = [vp, vp, vp, C.c_uint64, C.POINTER(ColorParams),
                                                   vp, vp, C.c_uint64, C.POINTER(C.c_uint32),
                                                   C.POINTER(C.c_float)]
        L.sbwt_index_build_color_table.argtypes = [vp, vp, C.POINTER(C.c_int)]
        L.sbwt_searcher_set_color_output.argtypes = [vp, C.c_int]
        L.sbwt_color_searcher_set_input.argtypes = [vp, vp]
        L.sbwt_color_searcher_set_chain_input.argtypes = [vp, vp]
        L.sbwt_color_searcher_last_time.argtypes = [vp, C.POINTER(C.c_float)]
        L.sbwt_color_searcher_set_census.argtypes = [vp, C.c_int]
        L.sbwt_searcher_last_key_runs_time.argtypes = [vp, C.POINTER(C.c_float)]
        L.sbwt_color_searcher_last_census.argtypes = [vp, u64p]
        L.sbwt_index_file_parser_create.argtypes = [C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_index_file_parser_destroy.argtypes = [vp]
        L.sbwt_index_file_parser_destroy.restype = None
        L.sbwt_index_file_parser_parse_device.argtypes = [vp, C.c_int, vp, C.c_uint64,
                                                          C.POINTER(vp), C.POINTER(vp), u64p, u64p]
        L.sbwt_index_file_parser_parse.argtypes = [vp, C.c_int, C.c_char_p, C.c_uint64, u64p,
                                                   C.c_uint64, u64p, C.c_uint64, u64p, u64p]
        L.sbwt_index_file_batch_cut.argtypes = [C.c_int, C.c_char_p, C.c_uint64, C.c_uint64]
        L.sbwt_index_file_batch_cut.restype = C.c_uint64
        L.sbwt_index_file_colored_cut.argtypes = [C.c_int, C.c_char_p, C.c_uint64]
        L.sbwt_index_file_colored_cut.restype = C.c_uint64
        L.sbwt_searcher_search_reads_format.argtypes = [vp, vp, C.c_int, C.c_char_p, C.c_uint64,
                                                        u64p, C.c_char_p, C.c_uint64, C.c_int,
                                                        C.c_char_p, C.c_uint64, u64p]
        L.sbwt_formatter_copy_last.argtypes = [vp, C.c_char_p, C.c_uint64]
        L.sbwt_fasta_parser_create.argtypes = [C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_fasta_parser_destroy.argtypes = [vp]
        L.sbwt_fasta_parser_destroy.restype = None
        L.sbwt_fasta_parser_parse.argtypes = [vp, C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint64,
                                              u64p, C.c_uint64, u64p, u64p]
        L.sbwt_fasta_chunk_cut.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.c_uint64]
        L.sbwt_fasta_chunk_cut.restype = C.c_uint64
        L.sbwt_fasta_has_quality_line.argtypes = [C.c_char_p, C.c_uint64, C.c_int]
        L.sbwt_searcher_search_fasta_format.argtypes = [vp, vp, vp, C.c_int, C.c_char_p,
                                                        C.c_uint64, C.c_char_p, C.c_uint64, u64p,
                                                        u64p]
        L.sbwt_fastq_parser_create.argtypes = [C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_fastq_parser_destroy.argtypes = [vp]
        L.sbwt_fastq_parser_destroy.restype = None
        L.sbwt_fastq_parser_parse.argtypes = [vp, C.c_char_p, C.c_uint64, C.c_int, C.c_char_p,
                                              C.c_uint64, u64p, C.c_uint64, u64p, u64p, u64p]
        L.sbwt_searcher_search_fastq_format.argtypes = [vp, vp, vp, C.c_int, C.c_char_p,
                                                        C.c_uint64, u64p]
        L.sbwt_color_formatter_create.argtypes = [C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]
        L.sbwt_color_formatter_destroy.argtypes = [vp]
        L.sbwt_color_formatter_destroy.restype = None
        L.sbwt_color_formatter_format.argtypes = [vp, C.c_int, vp, vp, C.c_uint64, C.c_uint64, vp,
                                                  C.c_uint64, u64p]
        L.sbwt_color_file_searcher_create.argtypes = [vp, C.c_uint64, C.c_uint64, vp,
                                                      C.POINTER(vp)]
        L.sbwt_color_file_searcher_destroy.argtypes = [vp]
        L.sbwt_color_file_searcher_destroy.restype = None
        L.sbwt_color_file_searcher_batch.argtypes = [vp, C.c_int, C.c_char_p, C.c_uint64,
                                                     C.POINTER(ColorParams), C.c_int, C.c_char_p,
                                                     C.c_uint64, u64p, u64p]
        L.sbwt_color_file_searcher_copy_last.argtypes = [vp, C.c_char_p, C.c_uint64]
        L.sbwt_server_create.argtypes = [vp, vp, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.sbwt_server_submit.argtypes = [vp, C.c_char_p, C.c_uint64, u64p, C.c_uint64, u64p]
        L.sbwt_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(ServerResult)]
        L.sbwt_server_release.argtypes = [vp, C.c_uint64]
        f64p = C.POINTER(C.c_double)
        L.sbwt_server_replay.argtypes = [vp, C.c_char_p, u64p, u64p, C.c_uint64, f64p, f64p,
                                         f64p, u64p, f64p]
        L.sbwt_server_get_slots.argtypes = [vp, C.POINTER(C.c_uint32), u64p]
        L.sbwt_server_destroy.argtypes = [vp]
        L.sbwt_server_destroy.restype = None
        L.sbwt_searcher_set_stats.argtypes = [vp, C.c_int]
        L.sbwt_searcher_last_stats.argtypes = [vp, u64p]
        L.sbwt_searcher_set_algorithm.argtypes = [vp, C.c_int]
        L.sbwt_searcher_set_launch.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.sbwt_searcher_get_launch.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.sbwt_searcher_autotune.argtypes = [vp, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                             C.c_void_p, C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        _lib = L
    return _lib


def header_symbols():
    """Names of every function declared in include/sbwt_amd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sbwt_[a-z0-9_]+)\s*\(", text)))


def _check(rc):
    if rc != 0:
        raise SbwtError(lib().sbwt_last_error().decode())


def _p64(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(u64p)


def mem_info():
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib().sbwt_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def suffix_groups_consistent(acgt, suffix_group_starts, num_bits):
    """Host-only check that every node outside the suffix-group starts has an
    empty row (sbwt_suffix_groups_consistent); no device work."""
    keep = [np.ascontiguousarray(a, np.uint64) for a in acgt]
    arr = (u64p * 4)(*[_p64(a) for a in keep])
    sg = np.ascontiguousarray(suffix_group_starts, np.uint64)
    rc = lib().sbwt_suffix_groups_consistent(arr, _p64(sg), num_bits)
    if rc < 0:
        raise SbwtError(lib().sbwt_last_error().decode())
    return bool(rc)


def random_line_probe(working_set_bytes, blocks_per_cu=8, steps=256):
    """Dependent random 128-byte line reads at a working-set size: returns
    {"ms", "lines", "table_bytes", "lines_per_s", "gbs"} (gbs counts 128 B/line)."""
    ms, lines, tb = C.c_float(), C.c_uint64(), C.c_uint64()
    _check(lib().sbwt_random_line_probe(working_set_bytes, blocks_per_cu, steps, C.byref(ms),
                                        C.byref(lines), C.byref(tb)))
    lps = lines.value / (ms.value / 1e3)
    return {"ms": ms.value, "lines": lines.value, "table_bytes": tb.value, "lines_per_s": lps,
            "gbs": lps * 128 / 1e9}


class Index:
    """Device-resident SBWT index (rank blocks + presearch table + key marks)."""

    def __init__(self, handle):
        self.h = handle
        info = IndexInfo()
        _check(lib().sbwt_index_get_info(self.h, C.byref(info)))
        self.num_nodes = info.num_nodes
        self.k = info.kmer_size
        self.presearch_letters = info.presearch_letters
        self.c_map = [info.c_map[i] for i in range(5)]
        self.device_bytes = info.device_bytes
        self.has_key_marks = bool(info.has_key_kmer_marks)
        self.wide = bool(info.wide)
        self.has_suffix_groups = bool(info.has_suffix_groups)
        self.suffix_groups_rejected = bool(info.suffix_groups_rejected)

    @classmethod
    def load(cls, sbwt_path, colors_path=None, presearch_letters=0):
        h = C.c_void_p()
        _check(lib().sbwt_index_load(os.fsencode(sbwt_path),
                                     os.fsencode(colors_path) if colors_path else None,
                                     presearch_letters, C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, acgt, num_bits, k, key_marks=None, presearch_letters=0,
               suffix_group_starts=None):
        keep = [np.ascontiguousarray(a, np.uint64) for a in acgt]
        arr = (u64p * 4)(*[_p64(a) for a in keep])
        km = np.ascontiguousarray(key_marks, np.uint64) if key_marks is not None else None
        sg = np.ascontiguousarray(suffix_group_starts, np.uint64) \
            if suffix_group_starts is not None else None
        h = C.c_void_p()
        _check(lib().sbwt_index_create(arr, num_bits, k, _p64(km) if km is not None else None,
                                       _p64(sg) if sg is not None else None,
                                       presearch_letters, C.byref(h)))
        return cls(h)

    def close(self):
        if self.h:
            lib().sbwt_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_color_table(self, colors):
        """Per-node key color table from a ColorIndex (fused path); returns its
        kind: 1 masks, 2 key ranks."""
        kind = C.c_int()
        _check(lib().sbwt_index_build_color_table(self.h, colors.h, C.byref(kind)))
        return kind.value

    def build_kmer_table(self):
        """Builds the k-mer table (fresh searches read a bucket pair instead of
        the presearch entry and SBWT steps); False when the index cannot have one."""
        built = C.c_int()
        _check(lib().sbwt_index_build_kmer_table(self.h, C.byref(built)))
        return bool(built.value)

    def searcher(self, max_chars, max_reads, move_to_key_kmer=False, stream=None):
        return Searcher(self, max_chars, max_reads, move_to_key_kmer, stream)

    def search_reads_noalloc(self, d_chars, n_chars, d_read_ends, n_reads, d_results,
                             move_to_key_kmer=False, stream=None):
        """Allocation-free fused search (sbwt_index_search_reads_noalloc): device
        buffers in and out, enqueued on `stream` (a hipStream_t, or None)."""
        _check(lib().sbwt_index_search_reads_noalloc(
            self.h, int(move_to_key_kmer), C.c_void_p(d_chars), n_chars,
            C.c_void_p(d_read_ends), n_reads, C.c_void_p(d_results),
            C.c_void_p(stream) if stream else None))


class Searcher:
    """One HIP stream + batch buffers (the reference's IndexSearcher)."""

    def __init__(self, index, max_chars, max_reads, move_to_key_kmer=False, stream=None):
        self.index = index
        self.h = C.c_void_p()
        _check(lib().sbwt_searcher_create(index.h, max_chars, max_reads, int(move_to_key_kmer),
                                          C.c_void_p(stream) if stream else None,
                                          C.byref(self.h)))

    def close(self):
        if self.h:
            lib().sbwt_searcher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def search(self, bit_seqs, positions):
        bit_seqs = np.ascontiguousarray(bit_seqs, np.uint64)
        positions = np.ascontiguousarray(positions, np.uint64)
        out = np.empty(len(positions), np.uint64)
        _check(lib().sbwt_searcher_search(self.h, _p64(bit_seqs), len(bit_seqs), _p64(positions),
                                          len(positions), _p64(out)))
        return out

    def search_reads(self, chars, read_ends):
        chars = bytes(chars)
        read_ends = np.ascontiguousarray(read_ends, np.uint64)
        out = np.empty(max(len(chars), 1), np.uint64)
        nk = C.c_uint64()
        _check(lib().sbwt_searcher_search_reads(self.h, chars, len(chars), _p64(read_ends),
                                                len(read_ends), _p64(out), C.byref(nk)))
        return out[:nk.value]

    def search_reads_device(self, d_chars, n_chars, d_read_ends, n_reads, d_results,
                            want_count=False):
        nk = C.c_uint64()
        _check(lib().sbwt_searcher_search_reads_device(
            self.h, C.c_void_p(d_chars), n_chars, C.c_void_p(d_read_ends), n_reads,
            C.c_void_p(d_results), C.byref(nk) if want_count else None))
        return nk.value if want_count else None

    def search_device(self, d_bit_seqs, d_positions, n, d_results):
        _check(lib().sbwt_searcher_search_device(self.h, C.c_void_p(d_bit_seqs),
                                                 C.c_void_p(d_positions), n,
                                                 C.c_void_p(d_results)))

    def kmer_ends_device(self):
        p = C.c_void_p()
        _check(lib().sbwt_searcher_kmer_ends_device(self.h, C.byref(p)))
        return p.value

    def set_stats(self, enable):
        _check(lib().sbwt_searcher_set_stats(self.h, int(enable)))

    def set_color_output(self, on):
        """COLOR_OUTPUT_ENTRIES (or True): write each found k-mer's color-table entry
        instead of its key node; COLOR_OUTPUT_CHAINS: the key chains, for a color
        searcher with chain input; COLOR_OUTPUT_NONE (or False): key nodes."""
        _check(lib().sbwt_searcher_set_color_output(self.h, int(on)))

    def last_stats(self):
        out = np.zeros(16, np.uint64)
        _check(lib().sbwt_searcher_last_stats(self.h, _p64(out)))
        return {"full_searches": int(out[0]), "steps": int(out[1]), "walk_steps": int(out[2]),
                "invalid": int(out[3]), "streamed": int(out[4]), "stream_extra_lines": int(out[5]),
                "stream_fallbacks": int(out[6]), "group_steps": int(out[7]),
                "absence_checks": int(out[8]), "absence_hits": int(out[9]),
                "absent_skipped": int(out[10]), "hash_lines": int(out[11])}

    @staticmethod
    def resolved(st):
        """K-mers a census accounts for: each is settled exactly one way."""
        return (st["full_searches"] + st["streamed"] - st["stream_fallbacks"] + st["invalid"]
                + st["absence_hits"] + st["absent_skipped"])

    @staticmethod
    def lines(st):
        """128-byte lines a census reads: presearch (or a first k-mer table pair),
        node lines and further k-mer table pairs."""
        return (st["full_searches"] + st["steps"] + st["streamed"] + st["stream_extra_lines"]
                + st["walk_steps"] + st["absence_checks"] + st.get("hash_lines", 0))

    def set_algorithm(self, algorithm):
        """ALGO_STREAMING, ALGO_PER_KMER (the reference's search shape), ALGO_AUTO (default)
        or ALGO_STREAMING_ADAPTIVE."""
        _check(lib().sbwt_searcher_set_algorithm(self.h, int(algorithm)))

    def set_launch(self, chunk_kmers, blocks_per_cu):
        _check(lib().sbwt_searcher_set_launch(self.h, int(chunk_kmers), int(blocks_per_cu)))

    def launch(self):
        ch, bpc = C.c_uint32(), C.c_uint32()
        _check(lib().sbwt_searcher_get_launch(self.h, C.byref(ch), C.byref(bpc)))
        return int(ch.value), int(bpc.value)

    def autotune(self, d_chars, n_chars, d_read_ends, n_reads, d_results):
        """Device pointers (ints). Returns (chunk_kmers, blocks_per_cu, search_ms)."""
        ch, bpc, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        _check(lib().sbwt_searcher_autotune(self.h, d_chars, n_chars, d_read_ends, n_reads,
                                            d_results, C.byref(ch), C.byref(bpc), C.byref(ms)))
        return int(ch.value), int(bpc.value), float(ms.value)

    def sync(self):
        _check(lib().sbwt_searcher_sync(self.h))

    def last_timings(self):
        a, b, c, d = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        _check(lib().sbwt_searcher_last_timings(self.h, C.byref(a), C.byref(b), C.byref(c)))
        _check(lib().sbwt_searcher_last_key_runs_time(self.h, C.byref(d)))
        return {"pack_ms": a.value, "offsets_ms": b.value, "search_ms": c.value,
                "key_runs_ms": d.value}


def pack_2bit_device(d_chars, n_chars, d_words, d_invalid, stream=None):
    _check(lib().sbwt_pack_2bit_device(C.c_void_p(d_chars), n_chars, C.c_void_p(d_words),
                                       C.c_void_p(d_invalid),
                                       C.c_void_p(stream) if stream else None))


FORMAT_ASCII, FORMAT_BINARY, FORMAT_BOOL = 0, 1, 2


class Formatter:
    """Device formatter of index results into the index output bytes."""

    def __init__(self, max_kmers, max_pieces, stream=None):
        self.h = C.c_void_p()
        _check(lib().sbwt_formatter_create(max_kmers, max_pieces, stream, C.byref(self.h)))

    def format_device(self, mode, d_results, d_piece_kmer_ends, d_piece_closes, n_pieces,
                      line_open, d_out=None, capacity=0):
        """Device pointers (ints). Returns the byte count (written if d_out fits it)."""
        n = C.c_uint64()
        _check(lib().sbwt_formatter_format(self.h, mode, d_results, d_piece_kmer_ends,
                                           d_piece_closes, n_pieces, int(line_open), d_out,
                                           capacity, C.byref(n)))
        return int(n.value)

    def format(self, mode, results, piece_kmer_ends, piece_closes, line_open=False):
        """Host arrays in, bytes out (via torch device buffers)."""
        import torch
        res = torch.from_numpy(np.ascontiguousarray(results, np.uint64).view(np.int64)).cuda()
        ke = torch.from_numpy(np.ascontiguousarray(piece_kmer_ends, np.uint64).view(np.int64)).cuda()
        cl = torch.from_numpy(np.ascontiguousarray(piece_closes, np.uint8)).cuda()
        n = self.format_device(mode, res.data_ptr(), ke.data_ptr(), cl.data_ptr(), len(ke),
                               line_open)
        out = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        n2 = self.format_device(mode, res.data_ptr(), ke.data_ptr(), cl.data_ptr(), len(ke),
                                line_open, out.data_ptr(), n)
        assert n2 == n
        return out[:n].cpu().numpy().tobytes()

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_formatter_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def search_reads_format(searcher, formatter, mode, chars, read_ends, read_closes=None,
                        line_open=False):
    """Host reads -> search -> formatted index output bytes, in one C call."""
    chars = bytes(chars)
    ends = np.ascontiguousarray(read_ends, np.uint64)
    closes = bytes(np.ones(len(ends), np.uint8) if read_closes is None
                   else np.ascontiguousarray(read_closes, np.uint8))
    n = C.c_uint64()
    _check(lib().sbwt_searcher_search_reads_format(searcher.h, formatter.h, mode, chars, len(chars),
                                                   _p64(ends), closes, len(ends), int(line_open),
                                                   None, 0, C.byref(n)))
    out = C.create_string_buffer(max(n.value, 1))
    _check(lib().sbwt_formatter_copy_last(formatter.h, out, n.value))
    return out.raw[:n.value]


def fasta_chunk_cut(raw, last, max_reads):
    """Longest prefix of FASTA bytes that a chunk may end with (0: none)."""
    return int(lib().sbwt_fasta_chunk_cut(bytes(raw), len(raw), int(last), max_reads))


def fasta_has_quality_line(raw, at_line_start=True):
    return bool(lib().sbwt_fasta_has_quality_line(bytes(raw), len(raw), int(at_line_start)))


class FastaParser:
    """Device parser of FASTA chunks (raw bytes -> chars + read ends)."""

    def __init__(self, max_bytes, max_reads, stream=None):
        self.h = C.c_void_p()
        _check(lib().sbwt_fasta_parser_create(max_bytes, max_reads,
                                              C.c_void_p(stream) if stream else None,
                                              C.byref(self.h)))
        self.max_bytes = max_bytes
        self.max_reads = max_reads

    def parse(self, raw):
        """Host bytes in -> (chars bytes, read_ends u64[])."""
        raw = bytes(raw)
        chars = C.create_string_buffer(max(len(raw), 1))
        ends = np.zeros(self.max_reads, np.uint64)
        nc, nr = C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_fasta_parser_parse(self.h, raw, len(raw), chars, len(raw), _p64(ends),
                                             len(ends), C.byref(nc), C.byref(nr)))
        return chars.raw[:nc.value], ends[:nr.value].copy()

    def search_format(self, searcher, formatter, mode, raw):
        """One chunk -> device parse -> search -> formatted bytes; (bytes, n_reads)."""
        raw = bytes(raw)
        n, nr = C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_searcher_search_fasta_format(searcher.h, self.h, formatter.h, mode, raw,
                                                       len(raw), None, 0, C.byref(n),
                                                       C.byref(nr)))
        out = C.create_string_buffer(max(n.value, 1))
        _check(lib().sbwt_formatter_copy_last(formatter.h, out, n.value))
        return out.raw[:n.value], int(nr.value)

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_fasta_parser_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class FastqParser:
    """Device parser of four-line FASTQ chunks (longest prefix of whole records
    that the record reader reads the same way)."""

    def __init__(self, max_bytes, max_reads, stream=None):
        self.h = C.c_void_p()
        _check(lib().sbwt_fastq_parser_create(max_bytes, max_reads,
                                              C.c_void_p(stream) if stream else None,
                                              C.byref(self.h)))
        self.max_bytes = max_bytes
        self.max_reads = max_reads

    def parse(self, raw, last=True):
        """Host bytes in -> (bytes used, chars bytes, read_ends u64[])."""
        raw = bytes(raw)
        chars = C.create_string_buffer(max(len(raw), 1))
        ends = np.zeros(self.max_reads, np.uint64)
        used, nc, nr = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_fastq_parser_parse(self.h, raw, len(raw), int(last), chars, len(raw),
                                             _p64(ends), len(ends), C.byref(used), C.byref(nc),
                                             C.byref(nr)))
        return int(used.value), chars.raw[:nc.value], ends[:nr.value].copy()

    def search_format(self, searcher, formatter, mode, raw, last=True):
        """One chunk -> device parse -> search -> formatted bytes; (bytes, bytes used, n_reads)."""
        raw = bytes(raw)
        used, nc, nr = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_fastq_parser_parse(self.h, raw, len(raw), int(last), None, 0, None, 0,
                                             C.byref(used), C.byref(nc), C.byref(nr)))
        n = C.c_uint64()
        _check(lib().sbwt_searcher_search_fastq_format(searcher.h, self.h, formatter.h, mode,
                                                       None, 0, C.byref(n)))
        out = C.create_string_buffer(max(n.value, 1))
        _check(lib().sbwt_formatter_copy_last(formatter.h, out, n.value))
        return out.raw[:n.value], int(used.value), int(nr.value)

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_fastq_parser_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ColorIndex:
    """Device-resident color index (.tcolors)."""

    def __init__(self, path):
        self.h = C.c_void_p()
        _check(lib().sbwt_color_index_load(os.fsencode(path), C.byref(self.h)))
        info = ColorIndexInfo()
        _check(lib().sbwt_color_index_get_info(self.h, C.byref(info)))
        self.num_colors = info.num_colors
        self.num_color_sets = info.num_color_sets
        self.num_nodes = info.num_nodes
        self.device_bytes = info.device_bytes

    def close(self):
        if self.h:
            lib().sbwt_color_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def searcher(self, max_indexes, max_reads, stream=None):
        return ColorSearcher(self, max_indexes, max_reads, stream)


def color_params(threshold=1.0, include_not_found=False, include_invalid=False,
                 first_read_quirk=False):
    return ColorParams(threshold, int(include_not_found), int(include_invalid),
                       int(first_read_quirk))


class ColorSearcher:
    """One HIP stream + batch buffers (the reference's ColorSearcher)."""

    def __init__(self, colors, max_indexes, max_reads, stream=None):
        self.colors = colors
        self.h = C.c_void_p()
        _check(lib().sbwt_color_searcher_create(colors.h, max_indexes, max_reads,
                                                C.c_void_p(stream) if stream else None,
                                                C.byref(self.h)))

    def close(self):
        if self.h:
            lib().sbwt_color_searcher_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def search(self, warped, warp_intervals):
        """Reference interface -> u64 counts [(len(warp_intervals) - 1), num_colors]."""
        warped = np.ascontiguousarray(warped, np.uint64)
        wi = np.ascontiguousarray(warp_intervals, np.uint64)
        n = max(len(wi) - 1, 0)
        out = np.zeros(max(n * self.colors.num_colors, 1), np.uint64)
        _check(lib().sbwt_color_searcher_search(self.h, _p64(warped), len(warped), _p64(wi),
                                                len(wi), _p64(out)))
        return out[:n * self.colors.num_colors].reshape(n, self.colors.num_colors)

    def pseudoalign(self, kmer_results, read_kmer_ends, params, ids_capacity=None):
        res = np.ascontiguousarray(kmer_results, np.uint64)
        ends = np.ascontiguousarray(read_kmer_ends, np.uint64)
        n = len(ends)
        cap = ids_capacity if ids_capacity is not None else max(n * self.colors.num_colors, 1)
        offsets = np.zeros(n + 1, np.uint64)
        ids = np.zeros(max(cap, 1), np.uint32)
        stats = np.zeros(max(3 * n, 1), np.uint32)
        nids = C.c_uint64()
        _check(lib().sbwt_color_searcher_pseudoalign(
            self.h, _p64(res) if len(res) else None, _p64(ends), n, C.byref(params),
            _p64(offsets), ids.ctypes.data_as(u32p), cap, stats.ctypes.data_as(u32p),
            C.byref(nids)))
        lists = [ids[int(offsets[r]):int(offsets[r + 1])].tolist() for r in range(n)]
        return lists, stats[:3 * n].reshape(n, 3), offsets, ids[:nids.value]

    def pseudoalign_device(self, d_results, d_kmer_ends, n_reads, params, d_offsets, d_ids,
                           ids_capacity, d_stats=None):
        nids = C.c_uint64()
        _check(lib().sbwt_color_searcher_pseudoalign_device(
            self.h, C.c_void_p(d_results), C.c_void_p(d_kmer_ends), n_reads, C.byref(params),
            C.c_void_p(d_offsets), C.c_void_p(d_ids), ids_capacity,
            C.c_void_p(d_stats) if d_stats else None, C.byref(nids)))
        return nids.value

    def set_input(self, index, chains=False):
        """None: results are key nodes; an Index: its color-table entries, or with
        chains, the key chains a searcher with COLOR_OUTPUT_CHAINS writes over it."""
        if chains:
            _check(lib().sbwt_color_searcher_set_chain_input(self.h, index.h))
        else:
            _check(lib().sbwt_color_searcher_set_input(self.h, index.h if index is not None else None))

    def set_work_items(self, item_kmers):
        """0: one wave per read; else a queue of (read, k-mer range) items."""
        _check(lib().sbwt_color_searcher_set_work_items(self.h, item_kmers))

    def autotune(self, d_results, d_kmer_ends, n_reads, params, d_offsets, d_ids, ids_capacity):
        """Times the work-item shapes on this batch and keeps the fastest: (item_kmers, ms)."""
        ik, ms = C.c_uint32(), C.c_float()
        _check(lib().sbwt_color_searcher_autotune(
            self.h, C.c_void_p(d_results), C.c_void_p(d_kmer_ends), n_reads, C.byref(params),
            C.c_void_p(d_offsets), C.c_void_p(d_ids), ids_capacity, C.byref(ik), C.byref(ms)))
        return int(ik.value), float(ms.value)

    def counts_device(self, d_results, d_kmer_ends, n_reads, d_counts):
        _check(lib().sbwt_color_searcher_counts_device(self.h, C.c_void_p(d_results),
                                                       C.c_void_p(d_kmer_ends), n_reads,
                                                       C.c_void_p(d_counts)))

    def last_time(self):
        ms = C.c_float()
        _check(lib().sbwt_color_searcher_last_time(self.h, C.byref(ms)))
        return ms.value

    def set_census(self, on):
        _check(lib().sbwt_color_searcher_set_census(self.h, 1 if on else 0))

    def last_census(self):
        """{set_lines, mark_lines, payload_lines, kmers} of the last call."""
        out = (C.c_uint64 * 4)()
        _check(lib().sbwt_color_searcher_last_census(self.h, out))
        return dict(zip(("set_lines", "mark_lines", "payload_lines", "kmers"), list(out)))


class Server:
    """Serving loop over resident indexes (sbwt_server_*)."""

    def __init__(self, index, colors=None, slots=4, max_chars_per_slot=1 << 25,
                 max_reads_per_slot=1 << 20, memory_fraction=0.9, move_to_key_kmer=False,
                 params=None, graphs=True, kmer_results=True):
        self.index, self.colors = index, colors
        cfg = ServerConfig(slots, max_chars_per_slot, max_reads_per_slot, memory_fraction,
                           int(move_to_key_kmer), params or color_params(), int(not graphs),
                           int(not kmer_results))
        self.h = C.c_void_p()
        _check(lib().sbwt_server_create(index.h, colors.h if colors else None, C.byref(cfg),
                                        C.byref(self.h)))
        n = C.c_uint32()
        _check(lib().sbwt_server_get_slots(self.h, C.byref(n), None))
        self.slots = n.value

    def submit(self, chars, read_ends):
        chars = bytes(chars)
        ends = np.ascontiguousarray(read_ends, np.uint64)
        rid = C.c_uint64()
        _check(lib().sbwt_server_submit(self.h, chars, len(chars), _p64(ends) if len(ends) else None,
                                        len(ends), C.byref(rid)))
        return rid.value

    def wait(self, rid):
        """-> dict(kmers, color_lists or None, submit_s, start_s, end_s); raises on failure."""
        r = ServerResult()
        rc = lib().sbwt_server_wait(self.h, rid, C.byref(r))
        try:
            if rc != 0:
                raise SbwtError(lib().sbwt_last_error().decode())
            kmers = np.ctypeslib.as_array(r.kmer_results, shape=(r.n_kmers,)).copy() \
                if r.n_kmers else np.zeros(0, np.uint64)
            lists = None
            if r.color_offsets:
                off = np.ctypeslib.as_array(r.color_offsets, shape=(r.n_reads + 1,))
                n_ids = int(off[-1])
                ids = np.ctypeslib.as_array(r.color_ids, shape=(n_ids,)).copy() if n_ids else \
                    np.zeros(0, np.uint32)
                lists = [ids[int(off[i]):int(off[i + 1])].tolist() for i in range(r.n_reads)]
            return {"kmers": kmers, "color_lists": lists, "submit_s": r.submit_s,
                    "start_s": r.start_s, "end_s": r.end_s}
        finally:
            lib().sbwt_server_release(self.h, rid)

    def replay(self, chars, read_ends, first_read, arrival_s):
        """Load generator (sbwt_server_replay): request i = reads
        [first_read[i], first_read[i+1]) of the concatenated batch, submitted at
        arrival_s[i] seconds; returns per-request latency (end - scheduled
        arrival), submit lag and output digest, and the wall time."""
        chars = np.ascontiguousarray(chars, np.uint8)
        ends = np.ascontiguousarray(read_ends, np.uint64)
        first = np.ascontiguousarray(first_read, np.uint64)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        n = len(arr)
        assert len(first) == n + 1
        lat = np.zeros(n, np.float64)
        lag = np.zeros(n, np.float64)
        dig = np.zeros(n, np.uint64)
        wall = C.c_double()
        f64p = C.POINTER(C.c_double)
        _check(lib().sbwt_server_replay(
            self.h, chars.ctypes.data_as(C.c_char_p), _p64(ends) if len(ends) else None,
            _p64(first), n, arr.ctypes.data_as(f64p), lat.ctypes.data_as(f64p),
            lag.ctypes.data_as(f64p), _p64(dig), C.byref(wall)))
        return {"latency_s": lat, "lag_s": lag, "digest": dig, "wall_s": wall.value}

    def peak_in_flight(self):
        n, p = C.c_uint32(), C.c_uint64()
        _check(lib().sbwt_server_get_slots(self.h, C.byref(n), C.byref(p)))
        return p.value

    def close(self):
        if self.h:
            lib().sbwt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


INDEX_FILE_ASCII, INDEX_FILE_BINARY = 0, 1
COLOR_FORMAT_ASCII, COLOR_FORMAT_BINARY, COLOR_FORMAT_CSV = 0, 1, 2


def index_file_batch_cut(fmt, body, max_reads):
    """Longest prefix of an index-file body that a batch may end with."""
    return int(lib().sbwt_index_file_batch_cut(fmt, bytes(body), len(body), max_reads))


def index_file_colored_cut(fmt, body):
    """Longest prefix ending after a read that has a found index (0 if none)."""
    return int(lib().sbwt_index_file_colored_cut(fmt, bytes(body), len(body)))


class IndexFileParser:
    """Device parser of index-result file bodies (the colors stage's input)."""

    def __init__(self, max_bytes, max_reads, stream=None):
        self.h = C.c_void_p()
        _check(lib().sbwt_index_file_parser_create(max_bytes, max_reads,
                                                   C.c_void_p(stream) if stream else None,
                                                   C.byref(self.h)))
        self.max_bytes = max_bytes
        self.max_reads = max_reads

    def parse(self, fmt, body):
        """Host bytes in -> (results u64[], read_kmer_ends u64[])."""
        body = bytes(body)
        res = np.zeros(len(body) // 2 + 1, np.uint64)
        ends = np.zeros(self.max_reads, np.uint64)
        nr, nl = C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_index_file_parser_parse(self.h, fmt, body, len(body), _p64(res), len(res),
                                                  _p64(ends), len(ends), C.byref(nr), C.byref(nl)))
        return res[:nr.value].copy(), ends[:nl.value].copy()

    def parse_device(self, fmt, d_body, n_bytes):
        """Device body (int pointer, 16-byte aligned) -> (d_results, d_ends, n_results, n_reads)."""
        dr, de = C.c_void_p(), C.c_void_p()
        nr, nl = C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_index_file_parser_parse_device(self.h, fmt, C.c_void_p(d_body), n_bytes,
                                                         C.byref(dr), C.byref(de), C.byref(nr),
                                                         C.byref(nl)))
        return dr.value, de.value, int(nr.value), int(nl.value)

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_index_file_parser_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ColorFormatter:
    """Device printer of color id lists (ascii / binary / csv bytes)."""

    def __init__(self, max_ids, max_reads, stream=None):
        self.h = C.c_void_p()
        _check(lib().sbwt_color_formatter_create(max_ids, max_reads,
                                                 C.c_void_p(stream) if stream else None,
                                                 C.byref(self.h)))

    def format(self, mode, offsets, ids, num_colors):
        """Host CSR in, bytes out (via torch device buffers)."""
        import torch
        off = torch.from_numpy(np.ascontiguousarray(offsets, np.uint64).view(np.int64)).cuda()
        idv = torch.from_numpy(np.ascontiguousarray(ids if len(ids) else [0], np.uint32)
                               .view(np.int32)).cuda()
        n_reads = len(offsets) - 1
        n = C.c_uint64()
        _check(lib().sbwt_color_formatter_format(self.h, mode, C.c_void_p(off.data_ptr()),
                                                 C.c_void_p(idv.data_ptr()), n_reads, num_colors,
                                                 None, 0, C.byref(n)))
        out = torch.empty(max(n.value, 1), dtype=torch.uint8, device="cuda")
        n2 = C.c_uint64()
        _check(lib().sbwt_color_formatter_format(self.h, mode, C.c_void_p(off.data_ptr()),
                                                 C.c_void_p(idv.data_ptr()), n_reads, num_colors,
                                                 C.c_void_p(out.data_ptr()), n.value,
                                                 C.byref(n2)))
        torch.cuda.synchronize()
        return out[:n.value].cpu().numpy().tobytes()

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_color_formatter_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ColorFileSearcher:
    """The colors command's device stage: index-file body -> color printer bytes."""

    def __init__(self, colors, max_bytes, max_reads, stream=None):
        self.colors = colors
        self.h = C.c_void_p()
        _check(lib().sbwt_color_file_searcher_create(colors.h, max_bytes, max_reads,
                                                     C.c_void_p(stream) if stream else None,
                                                     C.byref(self.h)))

    def batch(self, in_fmt, body, params, out_mode):
        """Returns (bytes, n_reads)."""
        body = bytes(body)
        cap = max(1024, 4 * len(body))
        out = C.create_string_buffer(cap)
        nb, nr = C.c_uint64(), C.c_uint64()
        _check(lib().sbwt_color_file_searcher_batch(self.h, in_fmt, body, len(body),
                                                    C.byref(params), out_mode, out, cap,
                                                    C.byref(nb), C.byref(nr)))
        if nb.value > cap:
            out = C.create_string_buffer(nb.value)
            _check(lib().sbwt_color_file_searcher_copy_last(self.h, out, nb.value))
        return out.raw[:nb.value], int(nr.value)

    def close(self):
        if getattr(self, "h", None):
            lib().sbwt_color_file_searcher_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
