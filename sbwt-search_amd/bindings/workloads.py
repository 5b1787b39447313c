"""Seeded synthetic workloads and index construction (ctypes over libsbwt_build.so).

Builds the inputs of the search path: genomes, reads, an SBWT (.sbwt plain-matrix
v0.1) and a color index (.tcolors sdsl-hybrid-v4). Nothing here is timed.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "builder", "libsbwt_build.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.bld_draw.argtypes = [C.c_uint64] * 3
        L.bld_draw.restype = C.c_uint64
        L.bld_genome_random.argtypes = [C.c_uint64, C.c_uint64, u8p]
        L.bld_genome_mutate.argtypes = [u8p, C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                        C.c_uint32, u8p, C.c_uint64]
        L.bld_genome_mutate.restype = C.c_uint64
        L.bld_reads.argtypes = [C.c_uint64, u8p, u64p, C.c_uint64, C.c_uint64, u32p, C.c_double,
                                C.c_double, C.c_double, C.c_int, C.c_char_p, u64p, u32p]
        L.bld_reads.restype = C.c_uint64
        L.bld_lognormal_lengths.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                            C.c_uint32, C.c_uint32, u32p]
        L.bld_uniform_lengths.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, u32p]
        L.bld_permutation.argtypes = [C.c_uint64, C.c_uint64, u64p]
        L.bld_sbwt_build.argtypes = [u8p, u64p, C.c_uint64, C.c_uint32, C.c_int, C.c_int]
        L.bld_sbwt_build.restype = C.c_void_p
        L.bld_sbwt_free.argtypes = [C.c_void_p]
        L.bld_sbwt_num_nodes.argtypes = [C.c_void_p]
        L.bld_sbwt_num_nodes.restype = C.c_uint64
        L.bld_sbwt_num_kmers.argtypes = [C.c_void_p]
        L.bld_sbwt_num_kmers.restype = C.c_uint64
        L.bld_sbwt_bits.argtypes = [C.c_void_p, C.c_int, u64p]
        L.bld_sbwt_write.argtypes = [C.c_void_p, C.c_char_p]
        L.bld_sbwt_lookup.argtypes = [C.c_void_p, u64p, C.c_uint64, u64p]
        L.bld_colors_build.argtypes = [C.c_void_p, u8p, u64p, C.c_uint64, u32p, C.c_uint32,
                                       C.c_int]
        L.bld_colors_build.restype = C.c_void_p
        L.bld_colors_free.argtypes = [C.c_void_p]
        for f in ("bld_colors_num_colors", "bld_colors_num_sets", "bld_colors_num_dense"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_uint64
        L.bld_colors_key_marks.argtypes = [C.c_void_p, u64p]
        L.bld_colors_write.argtypes = [C.c_void_p, C.c_char_p]
        L.bld_sbwt_lookup.restype = C.c_int
        L.bld_gpu_build.argtypes = [u8p, u64p, C.c_uint64, C.c_uint32, C.c_int, u32p, C.c_uint32,
                                    C.c_int, C.POINTER(C.c_void_p), C.POINTER(GpuBuildStats),
                                    C.c_char_p, C.c_uint64]
        L.bld_gpu_build.restype = C.c_void_p
        _lib = L
    return _lib


class GpuBuildStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("occurrences", "pairs", "kmers", "missing_prefixes",
                                           "dummies", "nodes",
                                           "sets", "key_nodes", "cycle_nodes",
                                           "peak_device_bytes")] + \
               [(f, C.c_double) for f in ("ms_keys", "ms_sort", "ms_sbwt", "ms_colors", "ms_total")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _p(a, t):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(t)


ACGT = np.frombuffer(b"ACGT", np.uint8)


def genome_random(seed, length):
    out = np.empty(length, np.uint8)
    lib().bld_genome_random(seed, length, _p(out, u8p))
    return out


def genome_mutate(root, seed, sub_rate, indel_rate=0.0, max_indel=0):
    cap = len(root) + len(root) // 4 + 1024
    out = np.empty(cap, np.uint8)
    n = lib().bld_genome_mutate(_p(root, u8p), len(root), seed, sub_rate, indel_rate, max_indel,
                                _p(out, u8p), cap)
    return out[:n].copy()


def concat(seqs):
    ends = np.cumsum([len(s) for s in seqs]).astype(np.uint64)
    return np.concatenate(seqs).astype(np.uint8, copy=False), ends


def reads(seed, genomes, genome_ends, lens, random_frac=0.0, err_rate=0.0, n_rate=0.0,
          alternate=False):
    """Returns (chars bytes-like np.uint8 array, ends u64, src_genome u32)."""
    lens = np.ascontiguousarray(lens, np.uint32)
    total = int(lens.sum(dtype=np.uint64))
    out = np.empty(total, np.uint8)
    ends = np.empty(len(lens), np.uint64)
    src = np.empty(len(lens), np.uint32)
    g = np.ascontiguousarray(genomes, np.uint8)
    ge = np.ascontiguousarray(genome_ends, np.uint64)
    lib().bld_reads(seed, _p(g, u8p), _p(ge, u64p), len(ge), len(lens), _p(lens, u32p),
                    random_frac, err_rate, n_rate, int(alternate),
                    out.ctypes.data_as(C.c_char_p), _p(ends, u64p), _p(src, u32p))
    return out, ends, src


def lognormal_lengths(seed, n, median, sigma, lo, hi):
    out = np.empty(n, np.uint32)
    lib().bld_lognormal_lengths(seed, n, median, sigma, lo, hi, _p(out, u32p))
    return out


def uniform_lengths(seed, n, lo, hi):
    out = np.empty(n, np.uint32)
    lib().bld_uniform_lengths(seed, n, lo, hi, _p(out, u32p))
    return out


def kmer_key(codes):
    """Colex key of a k-mer given as codes: char j at bits [2j, 2j+1]."""
    v = 0
    for j, c in enumerate(codes):
        v |= int(c) << (2 * j)
    return v


class BuiltSbwt:
    def __init__(self, codes, ends, k, add_rc=True, threads=0, handle=None):
        self.codes = np.ascontiguousarray(codes, np.uint8)
        self.ends = np.ascontiguousarray(ends, np.uint64)
        self.k = k
        if handle is None:
            handle = lib().bld_sbwt_build(_p(self.codes, u8p), _p(self.ends, u64p),
                                          len(self.ends), k, int(add_rc), threads)
        self.h = handle
        if not self.h:
            raise RuntimeError("SBWT construction failed")
        self.n = lib().bld_sbwt_num_nodes(self.h)
        self.n_kmers = lib().bld_sbwt_num_kmers(self.h)

    def bits(self, c):
        out = np.empty((self.n + 63) // 64, np.uint64)
        lib().bld_sbwt_bits(self.h, c, _p(out, u64p))
        return out

    def write(self, path):
        if lib().bld_sbwt_write(self.h, os.fsencode(path)):
            raise RuntimeError(f"writing {path} failed")

    def lookup(self, keys):
        keys = np.ascontiguousarray(keys, np.uint64)
        out = np.empty(len(keys), np.uint64)
        if lib().bld_sbwt_lookup(self.h, _p(keys, u64p), len(keys), _p(out, u64p)):
            raise RuntimeError("this SBWT was built without its k-mer list")
        return out

    def colors(self, color_of, d=20, threads=0):
        return BuiltColors(self, color_of, d, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bld_sbwt_free(self.h)
            self.h = None


class BuiltColors:
    def __init__(self, sbwt, color_of, d, threads, handle=None):
        self.sbwt = sbwt
        if handle is None:
            co = np.ascontiguousarray(color_of, np.uint32)
            handle = lib().bld_colors_build(sbwt.h, _p(sbwt.codes, u8p), _p(sbwt.ends, u64p),
                                            len(sbwt.ends), _p(co, u32p), d, threads)
        self.h = handle
        if not self.h:
            raise RuntimeError("color index construction failed")
        self.num_colors = lib().bld_colors_num_colors(self.h)
        self.num_sets = lib().bld_colors_num_sets(self.h)
        self.num_dense = lib().bld_colors_num_dense(self.h)

    def key_marks(self):
        out = np.empty((self.sbwt.n + 63) // 64, np.uint64)
        lib().bld_colors_key_marks(self.h, _p(out, u64p))
        return out

    def write(self, path):
        if lib().bld_colors_write(self.h, os.fsencode(path)):
            raise RuntimeError(f"writing {path} failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().bld_colors_free(self.h)
            self.h = None


def permutation(seed, n):
    out = np.empty(n, np.uint64)
    lib().bld_permutation(seed, n, _p(out, u64p))
    return out


def gpu_build(codes, ends, k, add_rc=True, color_of=None, d=20, keep_kmers=False):
    """The SBWT (and with color_of the color index) built on the current GPU:
    bit for bit the host builder's. Returns (BuiltSbwt, BuiltColors or None,
    stats dict)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    ends = np.ascontiguousarray(ends, np.uint64)
    co = None if color_of is None else np.ascontiguousarray(color_of, np.uint32)
    col_h = C.c_void_p()
    st = GpuBuildStats()
    err = C.create_string_buffer(512)
    h = lib().bld_gpu_build(_p(codes, u8p), _p(ends, u64p), len(ends), k, int(add_rc),
                            None if co is None else _p(co, u32p), d, int(keep_kmers),
                            C.byref(col_h), C.byref(st), err, len(err))
    if not h:
        raise RuntimeError(f"GPU construction failed: {err.value.decode()}; counts so far: "
                           f"{st.as_dict()}")
    sb = BuiltSbwt(codes, ends, k, add_rc, handle=h)
    colors = BuiltColors(sb, None, d, 0, handle=col_h.value) if co is not None else None
    return sb, colors, st.as_dict()
