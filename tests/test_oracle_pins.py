"""Pins the CPU oracle before it is trusted as the checker.

- index path end to end: byte-identical to the reference's own expected files
  (test_objects/full_pipeline/index_search/expected/*.txt);
- packing and positions: the reference's unit-test vectors;
- Poppy rank: the brute-force popcount rank on random and adversarial vectors;
- search: brute-force k-mer membership and colex rank on small built SBWTs,
  and invariance of the result under the presearch width.
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle_py as O
import workloads as W
from conftest import FIXTURES, GOLDEN

INDEX_FIXTURES = ["fasta1.fna", "fasta2.fna", "fasta3.fna", "fasta4.fna", "fastq1.fnq"]


@pytest.fixture(scope="module")
def fixture_sbwt():
    return O.OracleSbwt(os.path.join(FIXTURES, "search_test_index.sbwt"))


@pytest.mark.parametrize("name", INDEX_FIXTURES)
def test_oracle_ascii_matches_reference_expected(fixture_sbwt, name, tmp_path):
    out = tmp_path / "o.txt"
    fixture_sbwt.index_file(os.path.join(FIXTURES, "index_search", name), str(out), 0, True)
    expected = open(os.path.join(FIXTURES, "index_search", "expected",
                                 name.split(".")[0] + ".txt"), "rb").read()
    assert out.read_bytes() == expected


def _ascii_tokens(data):
    body = data[25:]  # u64 5 "ascii" u64 4 "v1.0"
    return [line.split() for line in body.decode().split("\n")[:-1]]


@pytest.mark.parametrize("name", INDEX_FIXTURES)
def test_oracle_binary_and_bool_agree_with_expected(fixture_sbwt, name, tmp_path):
    """Token-level equivalence of the three formats, as verify_index_results_equal.py checks."""
    src = os.path.join(FIXTURES, "index_search", name)
    exp = _ascii_tokens(open(os.path.join(FIXTURES, "index_search", "expected",
                                          name.split(".")[0] + ".txt"), "rb").read())
    fixture_sbwt.index_file(src, str(tmp_path / "b.bin"), 1, True)
    data = (tmp_path / "b.bin").read_bytes()
    assert data[:8] == struct.pack("<Q", 6) and data[8:14] == b"binary"
    vals = np.frombuffer(data[26:], np.uint64)
    lines, cur = [], []
    for v in vals:
        if v == O.NEWLINE:
            lines.append(cur)
            cur = []
        else:
            cur.append("-1" if v == O.NOT_FOUND else "-2" if v == O.INVALID else str(int(v)))
    assert lines == exp
    fixture_sbwt.index_file(src, str(tmp_path / "b.bool"), 2, True)
    data = (tmp_path / "b.bool").read_bytes()
    assert data[8:12] == b"bool" and data[20:24] == b"v2.0"
    blines = data[24:].decode().split("\n")[:-1]
    assert len(blines) == len(exp)
    for bl, el in zip(blines, exp):
        assert list(bl) == ["2" if t == "-2" else "1" if t == "-1" else "0" for t in el]


def test_fixture_index_finds_every_indexed_kmer(fixture_sbwt):
    chars, ends = O.read_seqfile(os.path.join(FIXTURES, "search_test_indexed.fna"))
    res = fixture_sbwt.index_reads(chars, ends, kp=12)
    assert len(res) > 0 and not np.any(res == O.NOT_FOUND) and not np.any(res == O.INVALID)


def test_packing_matches_reference_unit_vectors():
    vec = json.load(open(os.path.join(GOLDEN, "reference_unit_vectors.json")))
    for case in vec["packing"]:
        words, invalid = O.pack(case["seq"].encode())
        nw = len(case["words"])
        assert [int(w) for w in words[:nw]] == case["words"]
        assert list(np.nonzero(invalid)[0]) == case["invalid_positions"]


def test_packing_closed_form():
    rng = np.random.default_rng(3)
    chars = bytes(rng.choice(list(b"ACGTacgtNn.-"), 1000))
    words, invalid = O.pack(chars)
    code = {ord(c): i for i, c in enumerate("ACGT")}
    code.update({ord(c): i for i, c in enumerate("acgt")})
    for j, ch in enumerate(chars):
        got = (int(words[j // 32]) >> (62 - 2 * (j % 32))) & 3
        assert got == code.get(ch, 0)
        assert invalid[j] == (ch not in code)


def test_positions_match_reference_unit_vectors():
    vec = json.load(open(os.path.join(GOLDEN, "reference_unit_vectors.json")))
    for case in vec["positions"]:
        got = O.positions(np.array(case["read_ends"], np.uint64), case["k"])
        assert [int(x) for x in got] == case["expected"]


@pytest.mark.parametrize("num_bits", [1, 63, 64, 65, 255, 256, 257, 1000, 1023, 1024, 1025,
                                      2048, 5000, 65536])
@pytest.mark.parametrize("kind", ["random", "zeros", "ones", "sparse"])
def test_poppy_rank_equals_popcount_rank(num_bits, kind):
    rng = np.random.default_rng(num_bits)
    nw = (num_bits + 63) // 64
    if kind == "random":
        bits = rng.integers(0, 2**63, nw, dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, nw, dtype=np.uint64)
    elif kind == "zeros":
        bits = np.zeros(nw, np.uint64)
    elif kind == "ones":
        bits = np.full(nw, np.uint64(0xFFFFFFFFFFFFFFFF))
    else:
        bits = np.zeros(nw, np.uint64)
        for i in rng.integers(0, num_bits, 5):
            bits[i // 64] |= np.uint64(1) << np.uint64(i % 64)
    if num_bits % 64:
        bits[-1] &= np.uint64((1 << (num_bits % 64)) - 1)
    bits = np.concatenate([bits, np.zeros(8, np.uint64)])
    qs = sorted(set([0, num_bits] + list(range(0, num_bits + 1, max(1, num_bits // 300)))
                    + [q for q in (63, 64, 65, 255, 256, 257, 767, 768, 1023, 1024, 1025)
                       if q <= num_bits]))
    r, b, total = O.poppy_rank(bits, num_bits, qs)
    assert np.array_equal(r, b)
    assert total == int(b[-1])


def _tiny_sbwt(seed, glen, k):
    g = W.genome_random(seed, glen)
    codes, ends = W.concat([g])
    return g, W.BuiltSbwt(codes, ends, k)


def _node_strings(g, k):
    """Explicit node set of an SBWT with reverse complements: k-mers + dummies."""
    rc = lambda s: "".join("ACGT"[3 - "ACGT".index(c)] for c in reversed(s))
    s = "".join("ACGT"[c] for c in g)
    kmers = {s[i:i + k] for i in range(len(s) - k + 1)}
    kmers |= {rc(x) for x in kmers}
    suffixes = {x[1:] for x in kmers}
    nodes = set(kmers)
    nodes.add("$" * k)
    for x in kmers:
        if x[:-1] not in suffixes:
            for j in range(1, k):
                nodes.add("$" * (k - j) + x[:j])
    return sorted(nodes, key=lambda t: t[::-1].replace("$", "!"))


@pytest.mark.parametrize("seed,glen,k", [(1, 300, 5), (2, 2000, 9), (3, 500, 31)])
def test_search_equals_brute_force_colex_rank(seed, glen, k, tmp_path):
    g, b = _tiny_sbwt(seed, glen, k)
    nodes = _node_strings(g, k)
    assert len(nodes) == b.n
    path = str(tmp_path / "t.sbwt")
    b.write(path)
    s = O.OracleSbwt(path)
    rank = {x: i for i, x in enumerate(nodes)}
    rng = np.random.default_rng(seed)
    gs = "".join("ACGT"[c] for c in g)
    queries = [gs[i:i + k] for i in rng.integers(0, glen - k + 1, 200)]
    queries += ["".join(rng.choice(list("ACGT"), k)) for _ in range(300)]
    for q in queries:
        codes = np.array(["ACGT".index(c) for c in q], np.uint8)
        got = s.search_codes(codes)
        exp = rank.get(q, int(O.NOT_FOUND))
        assert got == exp, q


def test_search_k32_poly_runs_equals_brute_force(tmp_path):
    """k = 32, where every 64-bit value is a k-mer code: poly-T (2^64 - 1) and
    poly-A (0) k-mers, present and absent, against the explicit colex rank."""
    k = 32
    g = W.genome_random(8, 700)
    g = np.concatenate([g[:300], np.full(40, 3, np.uint8), g[300:]]).astype(np.uint8)
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, k)
    nodes = _node_strings(g, k)
    assert len(nodes) == b.n
    path = str(tmp_path / "t.sbwt")
    b.write(path)
    s = O.OracleSbwt(path)
    rank = {x: i for i, x in enumerate(nodes)}
    gs = "".join("ACGT"[c] for c in g)
    queries = ["T" * 32, "A" * 32, "C" * 32, "G" * 32, "T" * 31 + "A", "A" * 31 + "T"]
    queries += [gs[i:i + k] for i in range(0, len(gs) - k + 1, 7)]
    for q in queries:
        got = s.search_codes(np.array(["ACGT".index(c) for c in q], np.uint8))
        assert got == rank.get(q, int(O.NOT_FOUND)), q
    assert rank["T" * 32] != int(O.NOT_FOUND) and rank["A" * 32] != int(O.NOT_FOUND)
    # and through the reads path (packing, windows of 32 chars)
    reads = ("T" * 50 + "A" * 40 + "C" * 33 + gs[100:200]).encode()
    res = s.index_reads(reads, np.array([len(reads)], np.uint64))
    for i in range(len(reads) - k + 1):
        assert int(res[i]) == rank.get(reads[i:i + k].decode(), int(O.NOT_FOUND)), i


def test_presearch_width_does_not_change_results(tmp_path):
    g, b = _tiny_sbwt(7, 5000, 31)
    path = str(tmp_path / "t.sbwt")
    b.write(path)
    s = O.OracleSbwt(path)
    chars, ends, _ = W.reads(11, g, np.array([len(g)], np.uint64), np.full(200, 100, np.uint32),
                             random_frac=0.3, err_rate=0.02, n_rate=0.01)
    base = s.index_reads(chars.tobytes(), ends, kp=12)
    assert np.any(base < O.INVALID)
    for kp in (1, 4, 8, 10):
        assert np.array_equal(s.index_reads(chars.tobytes(), ends, kp=kp), base)


def test_sbwt_writer_round_trip(tmp_path):
    g, b = _tiny_sbwt(4, 3000, 31)
    path = str(tmp_path / "t.sbwt")
    b.write(path)
    s = O.OracleSbwt(path)
    assert s.n == b.n and s.k == 31
    for c in range(4):
        assert np.array_equal(s.bits(c)[: (b.n + 63) // 64], b.bits(c))
    assert np.array_equal(s.sgs()[: (b.n + 63) // 64], b.bits(4))
    cm = s.c_map()
    assert cm[0] == 1 and cm[4] == b.n


@pytest.mark.parametrize("seed,glen,k,kp", [(4, 400, 9, 5), (5, 3000, 15, 6), (6, 800, 31, 7)])
def test_empty_presearch_interval_iff_in_no_kmer(seed, glen, k, kp, tmp_path):
    """A kp-letter string has an empty SBWT interval exactly when no k-mer of the
    index (reverse complements included) contains it: the dummy nodes make every
    prefix of a k-mer end some node. This is what lets one presearch lookup settle
    every window that holds the string as not found."""
    g, b = _tiny_sbwt(seed, glen, k)
    path = str(tmp_path / "t.sbwt")
    b.write(path)
    left, right = O.OracleSbwt(path).presearch(kp)
    kmers = [x for x in _node_strings(g, k) if "$" not in x]
    inside = {x[i:i + kp] for x in kmers for i in range(k - kp + 1)}
    n_absent = 0
    for x in range(4 ** kp):
        s = "".join("ACGT"[(x >> (2 * (kp - 1 - i))) & 3] for i in range(kp))
        empty = int(left[x]) > int(right[x])
        assert empty == (s not in inside), s
        n_absent += empty
    assert 0 < n_absent < 4 ** kp


@pytest.mark.parametrize("name", INDEX_FIXTURES)
def test_ascii_helper_matches_reference_expected(fixture_sbwt, name):
    """The test-side ascii writer reproduces the reference's expected files from
    the oracle's results (so GPU tests may compare formatted bytes with it)."""
    chars, ends = O.read_seqfile(os.path.join(FIXTURES, "index_search", name))
    res = fixture_sbwt.index_reads(chars, ends)
    expected = open(os.path.join(FIXTURES, "index_search", "expected",
                                 name.split(".")[0] + ".txt"), "rb").read()
    assert O.ascii_index_output(res, ends, 30) == expected
