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
from colorlib import colored_seq_ids, ColoredIndex, kmer_ends_of, read_stats, threshold_lists, K


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
    counts = np.array([[0], [5]], np.uint64)
    stats = [(0, 5, 0), (6, 0, 0)]
    assert threshold_lists(counts, stats, 0.5, False, False, True) == [[], [0]]
    # ... unless not-found k-mers are included: (2^64 - 1) + 4 wraps to 3, minimum 2,
    # met by the next colored read's count 5 (quirk 2)
    assert threshold_lists(counts, stats, 0.5, True, False, True) == [[0], [0]]


def test_warp_layout_matches_reference_unit_vectors():
    """orc_warp_layout (IndexFileParser.cpp:36-74) against the expected vectors of the
    reference's own parser tests (BinaryIndexFileParser_test.cpp), including
    colored_seq_id: a read with no found index gets the slot of the next read that
    has one; the entry after the last read is the batch-end slot."""
    import json
    import os
    from conftest import GOLDEN
    vecs = json.load(open(os.path.join(GOLDEN, "reference_unit_vectors.json")))["warp_layout"]
    for v in vecs:
        reads = v["reads"]
        n = len(reads)
        res, ends = [], []
        for r in reads:
            res += [int(O.NOT_FOUND) if x == -1 else int(O.INVALID) if x == -2 else x for x in r]
            ends.append(len(res))
        res = np.array(res + [0], np.uint64)
        ends = np.array(ends, np.uint64)
        w = v["warp"]
        warped = np.zeros(len(res) + w * (n + 1), np.uint64)
        wi = np.zeros(n + 2, np.uint64)
        ni = np.zeros(1, np.uint64)
        cs = np.zeros(n + 1, np.uint64)
        f, nf, inv = (np.zeros(n + 1, np.uint64) for _ in range(3))
        nw = O.lib().orc_warp_layout(O.p64(res), O.p64(ends), n, w, O.p64(warped), O.p64(wi),
                                     O.p64(ni), O.p64(cs), O.p64(f), O.p64(nf), O.p64(inv))
        exp_cs = v["colored_seq_id"]
        assert [int(x) for x in cs[:n]] == exp_cs[:n], v["case"]
        # a whole-file vector holds one more entry: the open sequence after the last
        # read end, whose slot is the batch-end slot (number of colored reads); a
        # first-batch vector ends inside its last read, whose counts are partial
        whole = len(exp_cs) == n + 1
        if whole:
            assert int(ni[0]) - 1 == exp_cs[n], v["case"]
        m = n if whole else n - 1
        assert [int(x) for x in f[:m]] == v["found"][:m]
        assert [int(x) for x in nf[:m]] == v["not_found"][:m]
        assert [int(x) for x in inv[:m]] == v["invalid"][:m]
        if "warped" in v:
            exp_w = [int(O.NOT_FOUND) if x == -1 else x for x in v["warped"]]
            assert [int(x) for x in warped[:nw]] == exp_w
            assert [int(x) for x in wi[:int(ni[0])]] == v["warp_intervals"]


def test_quirk2_known_answers():
    """Reads with no found index are printed with the next colored read's counts
    against their own threshold (hand-derived answers, tau = 0.5)."""
    counts = np.array([[3, 1, 0], [0, 0, 0], [0, 0, 0], [0, 4, 2], [0, 0, 0]], np.uint64)
    stats = [(4, 0, 0), (0, 5, 0), (0, 0, 3), (4, 1, 0), (0, 2, 0)]
    # without include flags a blank read's threshold is 0: nothing printed
    assert threshold_lists(counts, stats, 0.5) == [[0], [], [], [1, 2], []]
    # with not-found included: read 1 (min 3) takes read 3's counts -> [1];
    # read 2 (min 0, invalid not included) prints nothing; read 3 (min 3) -> [1];
    # read 4 (min 1) has no colored read after it: batch end, nothing
    assert threshold_lists(counts, stats, 0.5, True) == [[0], [1], [], [1], []]
    # with both flags: read 2 (min 2) takes read 3's counts -> [1, 2]
    assert threshold_lists(counts, stats, 0.5, True, True) == [[0], [1], [1, 2], [1], []]
    assert colored_seq_ids(stats) == [0, 1, 1, 1, 2]


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


# ---------------------------------------------------------------------------
# Index-result file bodies (AsciiIndexFileParser.cpp:33-98,
# BinaryIndexFileParser.cpp:33-70) and color printers (Ascii/Binary/Csv
# ContinuousColorResultsPrinter.cpp).

def _ascii_body(results, kmer_ends):
    """Index results as the ascii printer writes them (pinned against the
    reference's expected index files by test_oracle_pins)."""
    import sys
    out = []
    st = 0
    for e in kmer_ends:
        toks = []
        for v in results[st:int(e)]:
            v = int(v)
            toks.append(b"-2" if v == int(O.INVALID) else b"-1" if v == int(O.NOT_FOUND) else
                        str(v).encode())
        out.append(b" ".join(toks) + b"\n")
        st = int(e)
    return b"".join(out)


def _binary_body(results, kmer_ends):
    words = []
    st = 0
    for e in kmer_ends:
        words += [int(v) for v in results[st:int(e)]] + [int(O.NEWLINE)]
        st = int(e)
    return np.asarray(words, np.uint64).tobytes()


def _random_results(seed, n_reads):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, n_reads)
    lens[rng.random(n_reads) < 0.2] = 0  # reads shorter than k: empty lines
    total = int(lens.sum())
    vals = rng.integers(0, 1 << 63, total, dtype=np.uint64) >> rng.integers(0, 63, total).astype(np.uint64)
    pick = rng.random(total)
    vals[pick < 0.15] = O.NOT_FOUND
    vals[(pick >= 0.15) & (pick < 0.25)] = O.INVALID
    vals[pick > 0.99] = np.uint64((1 << 64) - 4)  # the largest found index the formats allow
    return vals, np.cumsum(lens).astype(np.uint64)


@pytest.mark.parametrize("name", ["fasta1", "fasta2", "fasta3", "fasta4"])
def test_parse_index_body_on_reference_files(name):
    """The reference's own color-search inputs: as many reads as its expected colors
    files have lines, and per read the same found / not-found / invalid tokens."""
    import os
    from conftest import FIXTURES
    data = open(os.path.join(FIXTURES, "color_search", f"{name}.indexes.txt"), "rb").read()
    res, ends = O.parse_index_body(data[25:], 0)
    exp_lines = open(os.path.join(FIXTURES, "color_search", "expected", f"{name}.colors.txt"),
                     "rb").read()[25:].split(b"\n")[:-1]
    assert len(ends) == len(exp_lines)
    st = 0
    for line, e in zip(data[25:].split(b"\n")[:-1], ends):
        toks = line.split()
        got = [int(v) for v in res[st:int(e)]]
        want = [int(O.NOT_FOUND) if t == b"-1" else int(O.INVALID) if t == b"-2" else int(t)
                for t in toks]
        assert got == want
        st = int(e)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_parse_index_body_round_trip(seed):
    """Printer bytes -> parser -> the same results and read structure, ascii and
    binary (the two formats must agree)."""
    res, ke = _random_results(seed, 300)
    for fmt, body in ((0, _ascii_body(res, ke)), (1, _binary_body(res, ke))):
        r2, e2 = O.parse_index_body(body, fmt)
        assert np.array_equal(r2, res) and np.array_equal(e2, ke), fmt


def test_parse_index_body_known_answers():
    """Inputs outside what the printers write, answered by hand from the reference's
    char loop (AsciiIndexFileParser.cpp:46-63 with :85-98)."""
    NF, INV = int(O.NOT_FOUND), int(O.INVALID)
    cases = [
        # '-' takes the next char, skips digits, consumes one more char ('-' here)
        (b"-1-1 5-2\n--\n3\n-\n4 9\n", [NF, 1, 5, INV, 3, 9], [4, 4, 5, 6]),
        (b"a\n", [ord("a") - ord("0")], [1]),          # any char >= '0' is a digit
        (b"1\x802\n", [1, 2], [2]),                     # bytes >= 0x80 are negative chars
        (b"1 2\n3", [1, 2, 3], [2]),                    # a tail with no read end
        (b"-3 -12 -21\n", [NF, INV], [2]),              # only the char after '-' counts
        (b"\n\n7\n", [7], [0, 0, 1]),
        (b"-\n5\n", [], [0]),                           # '\n' after '-' is swallowed; 5 skipped
        (b"18446744073709551616 99\n", [0, 99], [2]),  # u64 arithmetic wraps
        (b"", [], []),
    ]
    for body, res, ends in cases:
        r, e = O.parse_index_body(body, 0)
        assert [int(x) for x in r] == res and [int(x) for x in e] == ends, body
    words = np.array([5, int(O.NEWLINE), NF, INV, int(O.NEWLINE), 7], np.uint64).tobytes()
    r, e = O.parse_index_body(words + b"\x01\x02", 1)  # a partial last word is not read
    assert [int(x) for x in r] == [5, NF, INV, 7] and [int(x) for x in e] == [1, 3]


@pytest.mark.parametrize("name", ["fasta1", "fasta2", "fasta3", "fasta4"])
def test_color_lists_format_on_reference_files(name):
    """Ascii color printer against the reference's expected colors files (lines
    compared without trailing spaces, as the reference's own verifier does)."""
    import os
    from conftest import FIXTURES
    exp = open(os.path.join(FIXTURES, "color_search", "expected", f"{name}.colors.txt"),
               "rb").read()[25:]
    lines = [[int(t) for t in ln.split()] for ln in exp.split(b"\n")[:-1]]
    offsets = np.cumsum([0] + [len(x) for x in lines]).astype(np.uint64)
    ids = np.array([c for x in lines for c in x] or [0], np.uint32)
    got = O.color_lists_format(offsets, ids[:int(offsets[-1])], 3, 0)
    assert [ln.rstrip(b" ") for ln in got.split(b"\n")] == \
        [ln.rstrip(b" ") for ln in exp.split(b"\n")]


def test_color_lists_format_binary_and_csv():
    """Binary: u64 ids then u64 max per read; csv: the indicator row of each read."""
    rng = np.random.default_rng(5)
    nc = 13
    lists = [sorted(rng.choice(nc, rng.integers(0, nc + 1), replace=False).tolist())
             for _ in range(50)]
    offsets = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
    ids = np.array([c for x in lists for c in x], np.uint32)
    words = np.frombuffer(O.color_lists_format(offsets, ids, nc, 1), np.uint64)
    back, cur = [], []
    for w in words:
        if w == np.uint64(~np.uint64(0)):
            back.append(cur)
            cur = []
        else:
            cur.append(int(w))
    assert back == lists
    rows = O.color_lists_format(offsets, ids, nc, 2).split(b"\n")[:-1]
    assert len(rows) == len(lists)
    for row, x in zip(rows, lists):
        assert row.split(b",") == [b"1" if c in x else b"0" for c in range(nc)]
    asc = O.color_lists_format(offsets, ids, nc, 0).split(b"\n")[:-1]
    assert [[int(t) for t in ln.split()] for ln in asc] == lists
