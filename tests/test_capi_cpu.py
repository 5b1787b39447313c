"""The C-ABI library loads on the CPU host and exports every declared symbol."""
import ctypes
import os
import subprocess

import pytest

import sbwt_amd as S


def test_library_exports_every_header_symbol():
    L = S.lib()
    names = S.header_symbols()
    assert "sbwt_index_create" in names and "sbwt_searcher_search" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "--notes", S.LIB_PATH],
                         capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("llvm-readelf unavailable")
    blob = open(S.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_last_error_starts_empty():
    assert S.lib().sbwt_last_error() in (b"",)


def test_load_rejects_wrong_sbwt_variant(tmp_path):
    """Format errors are reported before any device work (SbwtBuilder.cpp:64,67)."""
    import struct
    p = tmp_path / "bad.sbwt"
    p.write_bytes(struct.pack("<Q", 6) + b"nested" + struct.pack("<Q", 4) + b"v0.1")
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_load(os.fsencode(str(p)), None, 0, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == b"Error input is not a plain-matrix SBWT"
    p.write_bytes(struct.pack("<Q", 12) + b"plain-matrix" + struct.pack("<Q", 4) + b"v0.2")
    rc = S.lib().sbwt_index_load(os.fsencode(str(p)), None, 0, ctypes.byref(h))
    assert rc != 0 and S.lib().sbwt_last_error() == b"Error: wrong SBWT version"


def test_load_rejects_wrong_colors_magic(tmp_path):
    from conftest import FIXTURES
    p = tmp_path / "bad.tcolors"
    import struct
    p.write_bytes(struct.pack("<Q", 5) + b"hello")
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_load(os.fsencode(os.path.join(FIXTURES, "search_test_index.sbwt")),
                                 os.fsencode(str(p)), 0, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == \
        b"The colors file has an incorrect format. Expected 'sdsl-hybrid-v4'"


@pytest.mark.parametrize("kp,k", [(17, 31), (9, 8)])
def test_create_rejects_presearch_letters_out_of_range(kp, k):
    """The presearch prefix is at most min(k, 16) letters; checked before device work."""
    words = (ctypes.c_uint64 * 1)(0b1011)
    acgt = (ctypes.POINTER(ctypes.c_uint64) * 4)(*([ctypes.cast(words, ctypes.POINTER(ctypes.c_uint64))] * 4))
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_create(acgt, 4, k, None, None, kp, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == b"presearch letters must be in [1, min(k, 16)]"


def _fasta_bytes(seed, n):
    import numpy as np
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGT>@\n\n+x", np.uint8)
    b = alphabet[rng.integers(0, len(alphabet), n)]
    b[0] = ord(">")
    return b.tobytes()


@pytest.mark.parametrize("seed", range(6))
def test_fasta_chunk_cut_is_the_longest_record_prefix(seed):
    """Host-only cut rule, checked by brute force over its definition: the cut
    is the last header start (a '>'/'@' at a line start) before which at most
    max_reads records begin; at the end of the input the whole tail when it
    holds at most max_reads records."""
    raw = _fasta_bytes(seed, 3000 + seed * 977)
    n = len(raw)
    starts = [0] + [i for i in range(1, n) if raw[i - 1] == 10 and raw[i] in b">@"]
    for max_reads in (1, 2, 5, 40, 10_000):
        before = [h for k, h in enumerate(starts) if k <= max_reads and h > 0]
        want = max(before, default=0)
        assert S.fasta_chunk_cut(raw, False, max_reads) == want, max_reads
        want_last = n if len(starts) <= max_reads else want
        assert S.fasta_chunk_cut(raw, True, max_reads) == want_last, max_reads


def test_fasta_quality_line_detection():
    import re
    for seed in range(20):
        raw = _fasta_bytes(100 + seed, 50 + seed * 13)
        for at_start in (True, False):
            want = bool(re.search(rb"\n\+", raw)) or (at_start and raw[:1] == b"+")
            assert S.fasta_has_quality_line(raw, at_start) == want
    assert S.fasta_has_quality_line(b"+AC", True) and not S.fasta_has_quality_line(b"+AC", False)
    assert S.fasta_chunk_cut(b">a\nAC\n>b\nGT", False, 10) == 6
    assert S.fasta_chunk_cut(b">a\nACGT", False, 10) == 0
    assert S.fasta_chunk_cut(b">a\nACGT", True, 10) == 7


@pytest.mark.parametrize("seed", range(4))
def test_index_file_batch_cut_by_brute_force(seed):
    """Host-only cut rule of index-file batches, checked over its definition:
    ascii cuts right after the last '\\n' not preceded by '-' among the first
    max_reads '\\n' bytes; binary right after the last u64 max - 2 word among
    the first max_reads of them. Spans of 64 KiB are crossed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = 150_000 + seed * 33_333
    alphabet = np.frombuffer(b"0123456789 --\n", np.uint8)
    body = alphabet[rng.integers(0, len(alphabet), n)].tobytes()
    nls = [i for i in range(n) if body[i] == 10]
    for max_reads in (1, 2, 7, 1000, len(nls) - 1, len(nls), len(nls) + 5, 10 * n):
        limit = nls[max_reads - 1] + 1 if max_reads <= len(nls) else n
        want = max((i + 1 for i in nls if i + 1 <= limit and (i < 1 or body[i - 1] != ord("-"))),
                   default=0)
        assert S.index_file_batch_cut(0, body, max_reads) == want, max_reads
    words = rng.integers(0, 6, n // 8).astype(np.uint64)
    words[words == 5] = np.uint64(2**64 - 3)
    raw = words.tobytes() + b"\x01\x02\x03"  # a partial last word is not read
    ends = np.flatnonzero(words == np.uint64(2**64 - 3))
    for max_reads in (1, 3, 100, len(ends) - 1, len(ends), len(ends) + 1):
        want = 8 * (int(ends[min(max_reads, len(ends)) - 1]) + 1) if len(ends) else 0
        assert S.index_file_batch_cut(1, raw, max_reads) == want, max_reads
    assert S.index_file_batch_cut(0, b"12 -\n", 5) == 0
    assert S.index_file_batch_cut(0, b"12\n-\n", 5) == 3


def test_index_file_batch_cut_large_body():
    """Batches of 1 M units and more are counted in 64 parts first; the cut is
    the same rule: the max_reads-th read end is searched for around part
    boundaries (one thread's first and last part among them) and past the
    last one."""
    import numpy as np
    rng = np.random.default_rng(11)
    n = (12 << 20) + 5
    alphabet = np.frombuffer(b"0123456789 -\n", np.uint8)
    arr = alphabet[rng.integers(0, len(alphabet), n)]
    body = arr.tobytes()
    nls = np.flatnonzero(arr == 10)
    ok = nls[(nls < 1) | (arr[np.maximum(nls - 1, 0)] != ord("-"))]
    step = (n + 63) // 64
    probes = {1, 2, len(nls) - 1, len(nls), len(nls) + 1, 4 * n}
    for t in (1, 15, 16, 17, 40, 63):
        before = int(np.searchsorted(nls, t * step))  # '\n' bytes in the first t parts
        probes |= {before - 1, before, before + 1}
    for max_reads in sorted(p for p in probes if p >= 1):
        limit = int(nls[max_reads - 1]) + 1 if max_reads <= len(nls) else n
        below = ok[ok + 1 <= limit]
        want = int(below[-1]) + 1 if len(below) else 0
        assert S.index_file_batch_cut(0, body, max_reads) == want, max_reads
    words = rng.integers(0, 6, (1 << 21) + 3).astype(np.uint64)
    words[words == 5] = np.uint64(2**64 - 3)
    ends = np.flatnonzero(words == np.uint64(2**64 - 3))
    wstep = (len(words) + 63) // 64
    probes = {1, len(ends) - 1, len(ends), len(ends) + 1}
    for t in (1, 16, 31, 63):
        before = int(np.searchsorted(ends, t * wstep))
        probes |= {before - 1, before, before + 1}
    for max_reads in sorted(probes):
        want = 8 * (int(ends[min(max_reads, len(ends)) - 1]) + 1)
        assert S.index_file_batch_cut(1, words.tobytes(), max_reads) == want, max_reads


@pytest.mark.parametrize("seed", range(3))
def test_index_file_colored_cut_by_oracle_parse(seed):
    """The cut after the last read with a found index, checked against the oracle's
    parse of the same body (the reference's char loop): ascii lines of tokens
    including malformed ones ('-1-1' is a not-found then a found 1; '-\\n5' swallows
    the newline and the 5), and binary words."""
    import numpy as np
    import oracle_py as O
    rng = np.random.default_rng(seed)
    toks = [b"123", b"-1", b"-2", b"7", b"-1-1", b"-\n5", b"0", b"-12"]
    for trial in range(40):
        lines = []
        for _ in range(int(rng.integers(1, 12))):
            k = int(rng.integers(0, 5))
            pick = [toks[int(i)] for i in rng.integers(0, len(toks), k)]
            if rng.random() < 0.5:  # many reads with no found index
                pick = [t for t in pick if t in (b"-1", b"-2", b"-12")]
            lines.append(b" ".join(pick) + b"\n")
        body = b"".join(lines)
        res, kends = O.parse_index_body(body, 0)
        ends_b = [i + 1 for i in range(len(body)) if body[i] == 10 and (i < 1 or body[i - 1] != 45)]
        assert len(ends_b) == len(kends)
        want, st = 0, 0
        for r, e in enumerate(kends):
            if np.any(res[st:int(e)] < O.INVALID - np.uint64(1)):
                want = ends_b[r]
            st = int(e)
        assert S.index_file_colored_cut(0, body) == want, body
    nl, nf, inv = 2**64 - 3, 2**64 - 1, 2**64 - 2
    words = [5, nf, nl, nf, inv, nl, nl, 9, nl, nf, nl]
    raw = np.array(words, np.uint64).tobytes()
    assert S.index_file_colored_cut(1, raw) == 8 * 9
    assert S.index_file_colored_cut(1, raw[:8 * 8]) == 8 * 3
    assert S.index_file_colored_cut(1, np.array([nf, nl, nl], np.uint64).tobytes()) == 0
