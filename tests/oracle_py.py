"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only to check results or to time the CPU beside the GPU.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

NOT_FOUND = np.uint64(0xFFFFFFFFFFFFFFFF)
INVALID = np.uint64(0xFFFFFFFFFFFFFFFE)
NEWLINE = np.uint64(0xFFFFFFFFFFFFFFFD)
WALK_STUCK = np.uint64(0xFFFFFFFFFFFFFFFC)

u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)


class Poppy(C.Structure):
    _fields_ = [("layer_0", u64p), ("n_layer_0", C.c_uint64), ("layer_1_2", u64p),
                ("n_layer_1_2", C.c_uint64), ("total_1s", C.c_uint64)]


class Sbwt(C.Structure):
    _fields_ = [("num_bits", C.c_uint64), ("k", C.c_uint64), ("words", C.c_uint64),
                ("acgt", u64p * 4), ("poppy", Poppy * 4), ("c_map", C.c_uint64 * 5),
                ("sgs", u64p)]


class Intv(C.Structure):
    _fields_ = [("data", u64p), ("size", C.c_uint64), ("width", C.c_uint32)]


class Bitv(C.Structure):
    _fields_ = [("data", u64p), ("bits", C.c_uint64), ("poppy", Poppy)]


class Colors(C.Structure):
    _fields_ = [("dense_arrays", Bitv), ("dense_arrays_intervals", Intv),
                ("sparse_arrays", Intv), ("sparse_arrays_intervals", Intv),
                ("is_dense_marks", Bitv), ("key_kmer_marks", Bitv),
                ("color_set_idxs", Intv), ("num_color_sets", C.c_uint64),
                ("num_colors", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_sbwt_load.argtypes = [C.c_char_p, C.POINTER(Sbwt)]
        L.orc_sbwt_from_bits.argtypes = [u64p * 4, C.c_uint64, C.c_uint64, C.POINTER(Sbwt)]
        L.orc_sbwt_free.argtypes = [C.POINTER(Sbwt)]
        L.orc_poppy_build.argtypes = [u64p, C.c_uint64, C.POINTER(Poppy)]
        L.orc_poppy_free.argtypes = [C.POINTER(Poppy)]
        L.orc_rank.argtypes = [u64p, C.POINTER(Poppy), C.c_uint64]
        L.orc_rank.restype = C.c_uint64
        L.orc_rank_brute.argtypes = [u64p, C.c_uint64]
        L.orc_rank_brute.restype = C.c_uint64
        L.orc_presearch.argtypes = [C.POINTER(Sbwt), C.c_uint32, u64p, u64p]
        L.orc_search.argtypes = [C.POINTER(Sbwt), C.c_uint32, u64p, u64p, u64p, u64p, C.c_uint64,
                                 u64p, u64p, u8p, u32p, C.c_int]
        L.orc_search_codes.argtypes = [C.POINTER(Sbwt), u8p, C.c_uint64]
        L.orc_search_codes.restype = C.c_uint64
        L.orc_pack.argtypes = [C.c_char_p, C.c_uint64, u64p, u8p]
        L.orc_positions.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p]
        L.orc_positions.restype = C.c_uint64
        L.orc_index_reads.argtypes = [C.POINTER(Sbwt), C.c_uint32, u64p, u64p, C.c_char_p, u64p,
                                      C.c_uint64, u64p, u64p, C.c_int]
        L.orc_index_reads.restype = C.c_uint64
        L.orc_index_file.argtypes = [C.POINTER(Sbwt), C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                     u64p, C.c_int]
        L.orc_colors_load.argtypes = [C.c_char_p, C.POINTER(Colors)]
        L.orc_colors_free.argtypes = [C.POINTER(Colors)]
        L.orc_intv_get.argtypes = [C.POINTER(Intv), C.c_uint64]
        L.orc_intv_get.restype = C.c_uint64
        L.orc_color_set_of.argtypes = [C.POINTER(Colors), C.c_uint64, u32p]
        L.orc_color_set_of.restype = C.c_uint64
        L.orc_warp_layout.argtypes = [u64p, u64p, C.c_uint64, C.c_uint64, u64p, u64p, u64p, u64p,
                                      u64p, u64p, u64p]
        L.orc_warp_layout.restype = C.c_uint64
        L.orc_color_counts.argtypes = [C.POINTER(Colors), u64p, u64p, C.c_uint64, u64p, C.c_int]
        L.orc_color_min_hits.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_int,
                                         C.c_int]
        L.orc_color_min_hits.restype = C.c_uint64
        L.orc_parse_index_body.argtypes = [C.c_char_p, C.c_uint64, C.c_int, u64p, u64p, u64p]
        L.orc_parse_index_body.restype = C.c_uint64
        L.orc_color_lists_format.argtypes = [u64p, u32p, C.c_uint64, C.c_uint64, C.c_int,
                                             C.c_char_p]
        L.orc_color_lists_format.restype = C.c_uint64
        _lib = L
    return _lib


def p64(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(u64p)


def p8(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(u8p)


def p32(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(u32p)


class OracleSbwt:
    """An SBWT held by the oracle: loaded from a .sbwt file or from bit vectors."""

    def __init__(self, path=None, acgt=None, num_bits=None, k=None):
        self.s = Sbwt()
        if path is not None:
            rc = lib().orc_sbwt_load(os.fsencode(path), C.byref(self.s))
            if rc:
                raise RuntimeError(f"oracle could not load {path} (rc={rc})")
        else:
            self._keep = [np.ascontiguousarray(a, dtype=np.uint64) for a in acgt]
            arr = (u64p * 4)(*[p64(a) for a in self._keep])
            if lib().orc_sbwt_from_bits(arr, num_bits, k, C.byref(self.s)):
                raise RuntimeError("oracle could not build sbwt")
        self.n = self.s.num_bits
        self.k = self.s.k
        self._pre = {}

    def close(self):
        if self.s.num_bits:
            lib().orc_sbwt_free(C.byref(self.s))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bits(self, c):
        return np.ctypeslib.as_array(self.s.acgt[c], shape=(int(self.s.words),)).copy()

    def sgs(self):
        if not self.s.sgs:
            return None
        return np.ctypeslib.as_array(self.s.sgs, shape=(int(self.s.words),)).copy()

    def c_map(self):
        return [int(self.s.c_map[i]) for i in range(5)]

    def presearch(self, kp):
        if kp not in self._pre:
            n = 1 << (2 * kp)
            left = np.empty(n, np.uint64)
            right = np.empty(n, np.uint64)
            lib().orc_presearch(C.byref(self.s), kp, p64(left), p64(right))
            self._pre[kp] = (left, right)
        return self._pre[kp]

    def search(self, packed, positions, kp=12, key_marks=None, want_steps=False, threads=8):
        left, right = self.presearch(kp)
        packed = np.ascontiguousarray(packed, np.uint64)
        positions = np.ascontiguousarray(positions, np.uint64)
        out = np.empty(len(positions), np.uint64)
        steps = np.empty(len(positions), np.uint8) if want_steps else None
        walks = np.empty(len(positions), np.uint32) if want_steps else None
        km = p64(np.ascontiguousarray(key_marks, np.uint64)) if key_marks is not None else None
        lib().orc_search(C.byref(self.s), kp, p64(left), p64(right), p64(packed), p64(positions),
                         len(positions), km, p64(out), p8(steps) if want_steps else None,
                         p32(walks) if want_steps else None, threads)
        return (out, steps, walks) if want_steps else out

    def search_codes(self, codes):
        codes = np.ascontiguousarray(codes, np.uint8)
        return lib().orc_search_codes(C.byref(self.s), p8(codes), len(codes))

    def index_reads(self, chars, read_ends, kp=12, key_marks=None, threads=8):
        left, right = self.presearch(kp)
        read_ends = np.ascontiguousarray(read_ends, np.uint64)
        n_pos = lib().orc_positions(p64(read_ends), len(read_ends), self.k, None)
        out = np.empty(n_pos + 1, np.uint64)
        km = p64(np.ascontiguousarray(key_marks, np.uint64)) if key_marks is not None else None
        lib().orc_index_reads(C.byref(self.s), kp, p64(left), p64(right), bytes(chars),
                              p64(read_ends), len(read_ends), km, p64(out), threads)
        return out[:n_pos]

    def index_file(self, in_path, out_path, fmt=0, headers=True, key_marks=None, threads=4):
        km = p64(np.ascontiguousarray(key_marks, np.uint64)) if key_marks is not None else None
        rc = lib().orc_index_file(C.byref(self.s), os.fsencode(in_path), os.fsencode(out_path),
                                  fmt, int(headers), km, threads)
        if rc:
            raise RuntimeError(f"oracle index_file failed rc={rc}")


def poppy_rank(bits, num_bits, queries):
    """Reference-layout Poppy rank of each query (and the brute-force value)."""
    bits = np.ascontiguousarray(bits, np.uint64)
    P = Poppy()
    lib().orc_poppy_build(p64(bits), num_bits, C.byref(P))
    try:
        r = [lib().orc_rank(p64(bits), C.byref(P), int(q)) for q in queries]
        b = [lib().orc_rank_brute(p64(bits), int(q)) for q in queries]
        return np.array(r, np.uint64), np.array(b, np.uint64), int(P.total_1s)
    finally:
        lib().orc_poppy_free(C.byref(P))


def pack(chars):
    chars = bytes(chars)
    n = len(chars)
    words = np.zeros((n + 31) // 32 + 2, np.uint64)
    invalid = np.zeros(n + 1, np.uint8)
    lib().orc_pack(chars, n, p64(words), p8(invalid))
    return words, invalid[:n]


def positions(read_ends, k):
    read_ends = np.ascontiguousarray(read_ends, np.uint64)
    n = lib().orc_positions(p64(read_ends), len(read_ends), k, None)
    out = np.empty(n + 1, np.uint64)
    lib().orc_positions(p64(read_ends), len(read_ends), k, p64(out))
    return out[:n]


class Reads(C.Structure):
    _fields_ = [("chars", C.POINTER(C.c_char)), ("n_chars", C.c_uint64), ("ends", u64p),
                ("n_reads", C.c_uint64)]


def read_seqfile(path):
    """FASTA/FASTQ -> (bytes of all sequences concatenated, u64 cumulative read ends)."""
    L = lib()
    L.orc_read_seqfile.argtypes = [C.c_char_p, C.POINTER(Reads)]
    L.orc_reads_free.argtypes = [C.POINTER(Reads)]
    r = Reads()
    if L.orc_read_seqfile(os.fsencode(path), C.byref(r)):
        raise RuntimeError(f"cannot read {path}")
    try:
        chars = C.string_at(r.chars, r.n_chars)
        ends = np.ctypeslib.as_array(r.ends, shape=(int(r.n_reads),)).copy() if r.n_reads else \
            np.zeros(0, np.uint64)
        return chars, ends.astype(np.uint64)
    finally:
        L.orc_reads_free(C.byref(r))


def ascii_index_output(results, read_ends, k, headers=True):
    """Index results as the reference's ascii v1.0 file (test helper)."""
    import struct
    out = bytearray()
    if headers:
        out += struct.pack("<Q", 5) + b"ascii" + struct.pack("<Q", 4) + b"v1.0"
    i = 0
    start = 0
    for e in read_ends:
        e = int(e)
        nk = max(0, e - start - k + 1)
        toks = []
        for v in results[i:i + nk]:
            v = int(v)
            toks.append(b"-2" if v == int(INVALID) else b"-1" if v == int(NOT_FOUND) else
                        str(v).encode())
        out += b" ".join(toks) + b"\n"
        i += nk
        start = e
    return bytes(out)


def parse_index_body(body, fmt):
    """Oracle parse of an index-result file body (0 ascii, 1 binary) ->
    (results u64[], kmer_ends u64[] per read)."""
    n = len(body)
    res = np.zeros(n + 2, np.uint64)
    ends = np.zeros(n + 2, np.uint64)
    nres = np.zeros(1, np.uint64)
    nr = lib().orc_parse_index_body(bytes(body), n, fmt, p64(res), p64(nres), p64(ends))
    return res[:int(nres[0])].copy(), ends[:int(nr)].copy()


def color_lists_format(offsets, ids, num_colors, mode):
    """Oracle bytes of color id lists (0 ascii, 1 binary, 2 csv; no headers)."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    ids = np.ascontiguousarray(ids if len(ids) else np.zeros(1), np.uint32)
    n = len(offsets) - 1
    size = lib().orc_color_lists_format(p64(offsets), p32(ids), n, num_colors, mode, None)
    buf = C.create_string_buffer(max(int(size), 1))
    lib().orc_color_lists_format(p64(offsets), p32(ids), n, num_colors, mode, buf)
    return buf.raw[:int(size)]
