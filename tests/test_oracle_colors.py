"""Pins the oracle's color stage: .tcolors loading, color sets, counts, threshold.

The reference's color fixtures need a .tcolors file that is absent
(test_objects/themisto_example), so color parity is pinned by brute force on
self-generated indexes: the color set the oracle decodes for the key k-mer a
found k-mer walks to must equal the set of genomes holding that k-mer.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists, K


@pytest.fixture(scope="module")
def ci(tmp_path_factory):
    return ColoredIndex(tmp_path_factory.mktemp("col"), n_genomes=8, length=20_000, d=5)


def test_loader_sizes(ci):
    assert ci.num_colors == 8
    assert ci.oc.key_kmer_marks.bits == ci.sbwt.n
    assert ci.oc.num_color_sets == ci.colors.num_sets
    assert 0 < ci.colors.num_dense < ci.colors.num_sets  # both encodings occur


def test_color_set_of_walked_kmer_equals_brute_force(ci):
    kc = ci.kmer_colors()
    rng = np.random.default_rng(2)
    g = ci.genomes[3]
    for st in rng.integers(0, len(g) - K + 1, 300):
        x = g[st:st + K]
        chars = "".join("ACGT"[c] for c in x).encode()
        res = ci.index_results(chars, np.array([K], np.uint64))
        assert res[0] < O.INVALID
        import workloads as W
        assert ci.color_set(res[0]) == sorted(kc[W.kmer_key(x)])


def test_counts_equal_brute_force(ci):
    kc = ci.kmer_colors()
    chars, ends = ci.reads(120, 90, seed=4)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    counts = ci.oracle_counts(res, ke)
    import workloads as W
    st = 0
    for r, e in enumerate(ends):
        seq = np.frombuffer(chars[st:int(e)], np.uint8)
        exp = np.zeros(ci.num_colors, np.uint64)
        for p in range(len(seq) - K + 1):
            w = seq[p:p + K]
            if not all(ch in b"ACGT" for ch in w.tobytes()):
                continue
            codes = np.array([b"ACGT".index(ch) for ch in w.tobytes()], np.uint8)
            for c in kc.get(W.kmer_key(codes), ()):
                exp[c] += 1
        assert np.array_equal(counts[r], exp), r
        st = int(e)


@pytest.mark.parametrize("seq_size,tau", [(0, 0.7), (1, 0.7), (10, 0.7), (11, 0.7), (100, 0.7),
                                          (7, 1.0), (3, 0.0), (1000, 0.33), (2**53 + 1, 0.5)])
def test_threshold_formula(seq_size, tau):
    assert O.lib().orc_color_min_hits(seq_size, 0, 0, tau, 0, 0) == \
        int(math.ceil(float(seq_size) * tau))


def test_threshold_include_flags():
    L = O.lib()
    assert L.orc_color_min_hits(10, 5, 3, 0.5, 0, 0) == 5
    assert L.orc_color_min_hits(10, 5, 3, 0.5, 1, 0) == 8
    assert L.orc_color_min_hits(10, 5, 3, 0.5, 1, 1) == 9


def test_first_read_quirk_changes_threshold():
    """found 11, tau 0.7: ceil(7.7) = 8 normally, ceil(7.0) = 7 with the u64-max carry."""
    counts = np.array([[7, 8, 11]], np.uint64)
    assert threshold_lists(counts, [(11, 0, 0)], 0.7) == [[1, 2]]
    assert threshold_lists(counts, [(11, 0, 0)], 0.7, first_read_quirk=True) == [[0, 1, 2]]
    # a first read with no found k-mers wraps to 2^64 - 1 and reports nothing ...
    assert threshold_lists(np.array([[5]], np.uint64), [(0, 5, 0)], 0.5, False, False, True) == [[]]
    # ... unless not-found k-mers are included: (2^64 - 1) + 4 wraps to 3, minimum 2
    assert threshold_lists(np.array([[5]], np.uint64), [(0, 5, 0)], 0.5, True, False, True) == [[0]]


def _parse_ascii_indexes(data):
    """Restatement of AsciiIndexFileParser::generate_batch (AsciiIndexFileParser.cpp:33-98)
    over a whole file: per read, the found indexes and the not-found / invalid counts."""
    assert data[:8] == (5).to_bytes(8, "little") and data[8:13] == b"ascii"
    assert data[21:25] == b"v1.0"
    reads = []
    cur = {"found": [], "nf": 0, "inv": 0}
    for line_end, line in enumerate(data[25:].split(b"\n")[:-1]):
        cur = {"found": [], "nf": 0, "inv": 0}
        for tok in line.split():
            if tok == b"-1":
                cur["nf"] += 1
            elif tok == b"-2":
                cur["inv"] += 1
            else:
                cur["found"].append(int(tok))
        reads.append(cur)
    return reads


@pytest.mark.parametrize("name", ["fasta1", "fasta2", "fasta3", "fasta4"])
def test_warp_layout_on_reference_index_files(name):
    """The color stage input layout (IndexFileParser.cpp:36-74) on the reference's own
    color-search index files; the read count matches its expected colors files."""
    import os
    from conftest import FIXTURES
    data = open(os.path.join(FIXTURES, "color_search", f"{name}.indexes.txt"), "rb").read()
    reads = _parse_ascii_indexes(data)
    exp_lines = open(os.path.join(FIXTURES, "color_search", "expected", f"{name}.colors.txt"),
                     "rb").read()[25:].split(b"\n")[:-1]
    assert len(exp_lines) == len(reads)
    res, ends = [], []
    for r in reads:
        res += r["found"] + [int(O.NOT_FOUND)] * r["nf"] + [int(O.INVALID)] * r["inv"]
        ends.append(len(res))
    res = np.array(res, np.uint64)
    ends = np.array(ends, np.uint64)
    n = len(reads)
    warped = np.zeros(len(res) + 64 * n + 64, np.uint64)
    wi = np.zeros(n + 2, np.uint64)
    ni = np.zeros(1, np.uint64)
    cs = np.zeros(n + 1, np.uint64)
    f, nf, inv = (np.zeros(n + 1, np.uint64) for _ in range(3))
    nw = O.lib().orc_warp_layout(O.p64(res), O.p64(ends), n, 64, O.p64(warped), O.p64(wi),
                                 O.p64(ni), O.p64(cs), O.p64(f), O.p64(nf), O.p64(inv))
    assert nw % 64 == 0
    colored = [i for i, r in enumerate(reads) if r["found"]]
    assert int(ni[0]) == len(colored) + 1
    for q, i in enumerate(colored):
        assert int(cs[i]) == q
        span = int(wi[q + 1] - wi[q])
        assert span == (len(reads[i]["found"]) + 63) // 64
        blk = warped[int(wi[q]) * 64:int(wi[q + 1]) * 64]
        assert list(blk[:len(reads[i]["found"])]) == reads[i]["found"]
        assert np.all(blk[len(reads[i]["found"]):] == O.NOT_FOUND)
    for i, r in enumerate(reads):
        assert (int(f[i]), int(nf[i]), int(inv[i])) == (len(r["found"]), r["nf"], r["inv"])
