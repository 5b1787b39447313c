"""Index search on the MI355X through the C ABI, bit-exact against the oracle.

Every test prints the case it runs before touching the GPU.
"""
import os

import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W
from conftest import FIXTURES

pytestmark = pytest.mark.gpu

INDEX_FIXTURES = ["fasta1.fna", "fasta2.fna", "fasta3.fna", "fasta4.fna", "fastq1.fnq"]


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def fixture_pair():
    path = os.path.join(FIXTURES, "search_test_index.sbwt")
    return S.Index.load(path), O.OracleSbwt(path)


@pytest.fixture(scope="module")
def fixture_table_index():
    """The fixture index with its k-mer table (fresh searches read buckets)."""
    idx = S.Index.load(os.path.join(FIXTURES, "search_test_index.sbwt"))
    assert idx.build_kmer_table()
    return idx


FIXTURE_ALGOS = [(S.ALGO_AUTO, "auto"), (S.ALGO_STREAMING, "streaming"),
                 (S.ALGO_PER_KMER, "per-kmer")]


@pytest.mark.parametrize("table", [False, True], ids=["no-table", "kmer-table"])
@pytest.mark.parametrize("algo", [a for a, _ in FIXTURE_ALGOS], ids=[n for _, n in FIXTURE_ALGOS])
@pytest.mark.parametrize("name", INDEX_FIXTURES)
def test_reference_fixtures_byte_identical(fixture_pair, fixture_table_index, name, algo, table):
    """The reference's own index and expected files, through every search kernel:
    the streaming kernel (the bench's and the large-batch kernel) included."""
    say(f"fixture {name} algo={algo} table={table}")
    idx, orc = fixture_pair
    if table:
        idx = fixture_table_index
    chars, ends = O.read_seqfile(os.path.join(FIXTURES, "index_search", name))
    s = idx.searcher(max(len(chars), 1), max(len(ends), 1))
    s.set_algorithm(algo)
    s.set_stats(True)
    res = s.search_reads(chars, ends)
    assert np.array_equal(res, orc.index_reads(chars, ends))
    out = O.ascii_index_output(res, ends, idx.k)
    exp = open(os.path.join(FIXTURES, "index_search", "expected",
                            name.split(".")[0] + ".txt"), "rb").read()
    assert out == exp
    st = s.last_stats()
    if algo == S.ALGO_STREAMING:
        assert S.Searcher.resolved(st) == len(res)
        if np.count_nonzero(res[1:] < O.INVALID) and len(res) > 1:
            # consecutive found k-mers step from their predecessor's node
            both = (res[1:] < O.INVALID) & (res[:-1] < O.INVALID)
            assert st["streamed"] > 0 or not both.any()


@pytest.mark.parametrize("table", [False, True], ids=["no-table", "kmer-table"])
@pytest.mark.parametrize("algo,launch", [(S.ALGO_STREAMING_ADAPTIVE, None),
                                         (S.ALGO_STREAMING, (64, 2))],
                         ids=["adaptive", "streaming-64"])
def test_reference_fixtures_repeated_streaming(fixture_pair, fixture_table_index, algo, launch,
                                               table):
    """Every fixture's reads repeated past 2^16 chars, so that the adaptive form
    streams too (with work items shrunk to 8+ k-mers); results are the tiled
    oracle results and the reference's expected file repeated."""
    idx, orc = fixture_pair
    if table:
        idx = fixture_table_index
    for name in INDEX_FIXTURES:
        say(f"fixture {name} repeated algo={algo} launch={launch} table={table}")
        chars, ends = O.read_seqfile(os.path.join(FIXTURES, "index_search", name))
        if len(chars) == 0:  # a file of empty reads has nothing to repeat
            continue
        reps = (1 << 16) // len(chars) + 2
        big = bytes(chars) * reps
        big_ends = np.concatenate([ends + np.uint64(i * len(chars)) for i in range(reps)])
        assert len(big) >= (1 << 16)
        s = idx.searcher(len(big), len(big_ends))
        s.set_algorithm(algo)
        if launch:
            s.set_launch(*launch)
        s.set_stats(True)
        res = s.search_reads(big, big_ends)
        one = orc.index_reads(chars, ends)
        assert np.array_equal(res, np.tile(one, reps))
        st = s.last_stats()
        assert S.Searcher.resolved(st) == len(res)
        if np.count_nonzero(one < O.INVALID) > 1:
            assert st["streamed"] + st["full_searches"] > 0
        exp = open(os.path.join(FIXTURES, "index_search", "expected",
                                name.split(".")[0] + ".txt"), "rb").read()
        body = O.ascii_index_output(res, big_ends, idx.k, headers=False)
        one_body = O.ascii_index_output(one, ends, idx.k, headers=False)
        assert body == one_body * reps
        assert exp.endswith(one_body)


def test_rejected_suffix_groups_search_per_kmer(tmp_path):
    """A file whose rows break the SBWT property (an edge on a node that does not
    start its suffix group) is searched per k-mer whatever the algorithm: the
    results are the reference's interval search on the file's rows (the oracle)."""
    from sbwt_file import PlainMatrix
    say("rejected suffix-group starts: reference fixture with a start cleared")
    pm = PlainMatrix(os.path.join(FIXTURES, "search_test_index.sbwt"))
    edges = pm.vector(0) | pm.vector(1) | pm.vector(2) | pm.vector(3)
    node = next(i for i in range(1, pm.num_bits) if (int(edges[i // 64]) >> (i % 64)) & 1)
    pm.set_bit(4, node, 0)
    bad = str(tmp_path / "bad_starts.sbwt")
    pm.write(bad)
    orc = O.OracleSbwt(bad)
    idx = S.Index.load(bad)
    assert idx.suffix_groups_rejected and not idx.has_suffix_groups
    for name in INDEX_FIXTURES:
        chars, ends = O.read_seqfile(os.path.join(FIXTURES, "index_search", name))
        s = idx.searcher(max(len(chars), 1), max(len(ends), 1))
        s.set_algorithm(S.ALGO_STREAMING)
        s.set_stats(True)
        res = s.search_reads(chars, ends)
        assert np.array_equal(res, orc.index_reads(chars, ends))
        assert s.last_stats()["streamed"] == 0
    say("rejected suffix-group starts: builder index with an edge on a non-start row")
    root = W.genome_random(12, 20_000)
    gens = [root] + [W.genome_mutate(root, s_, 0.03) for s_ in (1, 2, 3)]
    codes, gends = W.concat(gens)
    b = W.BuiltSbwt(codes, gends, 31)
    good = str(tmp_path / "good.sbwt")
    b.write(good)
    pm = PlainMatrix(good)
    starts = pm.vector(4)
    non_start = [i for i in range(1, b.n) if not (int(starts[i // 64]) >> (i % 64)) & 1]
    for v in non_start[::max(1, len(non_start) // 50)]:
        pm.set_bit(v % 4, v, 1)
    bad2 = str(tmp_path / "bad_rows.sbwt")
    pm.write(bad2)
    assert S.Index.load(good).has_suffix_groups
    idx2 = S.Index.load(bad2)
    assert idx2.suffix_groups_rejected
    orc2 = O.OracleSbwt(bad2)
    chars, rends, _ = W.reads(7, codes, gends, np.full(2000, 150, np.uint32), 0.05, 0.005, 0.0)
    chars = chars.tobytes()
    for algo in (S.ALGO_STREAMING, S.ALGO_STREAMING_ADAPTIVE, S.ALGO_AUTO):
        s = idx2.searcher(len(chars), len(rends))
        s.set_algorithm(algo)
        assert np.array_equal(s.search_reads(chars, rends), orc2.index_reads(chars, rends)), algo


@pytest.mark.parametrize("algo", [S.ALGO_STREAMING, S.ALGO_PER_KMER, S.ALGO_STREAMING_ADAPTIVE],
                         ids=["streaming", "per-kmer", "adaptive"])
@pytest.mark.parametrize("present", [True, False], ids=["polyT-present", "polyT-absent"])
def test_kmer_table_k32_poly_t(tmp_path, present, algo):
    """k = 32 with the k-mer table: poly-T's code is 2^64 - 1 and poly-A's is 0,
    so bucket occupancy cannot be a code value. Reads full of T and A runs are
    searched with the index holding those k-mers and without them."""
    say(f"k=32 k-mer table poly-T present={present} algo={algo}")
    g = W.genome_random(321, 30_000)
    if present:
        g = np.concatenate([g[:10_000], np.full(40, 3, np.uint8), g[10_000:20_000],
                            np.full(35, 0, np.uint8), g[20_000:]])
    else:  # break every run of T or A longer than 20
        for c in (0, 3):
            run = 0
            for i in range(len(g)):
                run = run + 1 if g[i] == c else 0
                if run > 20:
                    g[i] = 1
                    run = 0
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, 32)
    path = str(tmp_path / "k32.sbwt")
    b.write(path)
    orc = O.OracleSbwt(path)
    idx = S.Index.load(path)
    assert idx.k == 32 and idx.build_kmer_table()
    rd, rends, _ = W.reads(9, codes, ends, np.full(3000, 120, np.uint32), 0.1, 0.01, 0.0)
    polys = [b"T" * 120, b"A" * 100, b"C" * 10 + b"T" * 60 + b"G" * 50, b"T" * 32, b"A" * 33,
             b"T" * 31 + b"A" * 40]
    chars = rd.tobytes() + b"".join(polys)
    rends = np.concatenate([rends, rends[-1] + np.cumsum([len(p) for p in polys]).astype(np.uint64)])
    s = idx.searcher(len(chars), len(rends))
    s.set_algorithm(algo)
    res = s.search_reads(chars, rends)
    exp = orc.index_reads(chars, rends)
    assert np.array_equal(res, exp)
    tail = exp[-(sum(max(0, len(p) - 31) for p in polys)):]
    if present:
        assert np.count_nonzero(tail < O.INVALID) > 0
    else:
        assert np.all(tail == O.NOT_FOUND)


@pytest.fixture(scope="module")
def genome_index(tmp_path_factory):
    g = W.genome_random(0, 200_000)
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(tmp_path_factory.mktemp("idx") / "g.sbwt")
    b.write(path)
    return g, b, path, O.OracleSbwt(path)


def _reads(g, n, length, seed, random_frac=0.5, err=0.01, n_rate=0.001):
    lens = np.full(n, length, np.uint32) if np.isscalar(length) else np.asarray(length, np.uint32)
    chars, ends, _ = W.reads(seed, g, np.array([len(g)], np.uint64), lens, random_frac, err,
                             n_rate)
    return chars.tobytes(), ends


ALGOS = [S.ALGO_STREAMING, S.ALGO_PER_KMER]


@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("kp", [12, 1, 7, 10, 13, 16])
def test_reads_parity_presearch_widths(genome_index, kp, algo):
    g, b, path, orc = genome_index
    say(f"reads parity kp={kp} algo={algo}")
    idx = S.Index.load(path, presearch_letters=kp)
    chars, ends = _reads(g, 3000, 150, seed=kp)
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(algo)
    res = s.search_reads(chars, ends)
    exp = orc.index_reads(chars, ends, kp=12)
    assert len(res) == len(exp) == 3000 * 120
    assert np.array_equal(res, exp)
    found = np.count_nonzero(res < O.INVALID)
    assert found > 0.2 * len(res)


def test_positions_interface_parity(genome_index):
    """The reference's IndexSearcher::search interface: bit_seqs + positions."""
    g, b, path, orc = genome_index
    say("positions interface")
    idx = S.Index.load(path)
    chars, ends = _reads(g, 500, 150, seed=3)
    packed, _ = O.pack(chars)
    pos = O.positions(ends, 31)
    rng = np.random.default_rng(1)
    pos = np.ascontiguousarray(rng.permutation(pos))
    s = idx.searcher(len(chars), len(ends))
    res = s.search(packed, pos)
    assert np.array_equal(res, orc.search(packed, pos, kp=12))
    with pytest.raises(S.SbwtError):
        s.search(packed[:2], np.array([1000], np.uint64))


@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
def test_ragged_reads_and_edge_cases(genome_index, algo):
    g, b, path, orc = genome_index
    say(f"ragged reads algo={algo}")
    idx = S.Index.load(path)
    lens = [0, 1, 30, 31, 32, 33, 63, 64, 65, 200, 1000, 5000, 0, 31]
    chars, ends = _reads(g, len(lens), lens, seed=5, random_frac=0.3, n_rate=0.01)
    chars = bytearray(chars)
    # lower case, all-N and high bytes inside some reads
    st = int(ends[8])
    chars[st:st + 20] = chars[st:st + 20].lower()
    st = int(ends[9])
    chars[st + 100:st + 140] = b"N" * 40
    chars[st + 500] = 0xC1
    chars = bytes(chars)
    s = idx.searcher(len(chars) + 100, len(ends) + 10)
    s.set_algorithm(algo)
    res = s.search_reads(chars, ends)
    assert np.array_equal(res, orc.index_reads(chars, ends))
    say("empty batch")
    assert len(s.search_reads(b"", np.zeros(0, np.uint64))) == 0
    say("reads shorter than k around a longer one")
    short = b"ACGT" * 10 + b"ACGTA"
    ends2 = np.array([5, 10, 45], np.uint64)  # 5, 5 and 35 chars: 0, 0 and 5 k-mers
    r2 = s.search_reads(short, ends2)
    assert len(r2) == 5 and np.array_equal(r2, orc.index_reads(short, ends2))
    say("reads all shorter than k")
    assert len(s.search_reads(b"ACGT" * 10, np.array([5, 10, 40], np.uint64))) == 0


def test_searcher_reuse_shrinking_batches(genome_index):
    g, b, path, orc = genome_index
    say("searcher reuse")
    idx = S.Index.load(path)
    s = idx.searcher(2_000_000, 20_000)
    for n, seed in [(5000, 1), (10, 2), (1200, 3)]:
        chars, ends = _reads(g, n, 150, seed=seed)
        assert np.array_equal(s.search_reads(chars, ends), orc.index_reads(chars, ends))
    with pytest.raises(S.SbwtError):
        chars, ends = _reads(g, 30_000, 150, seed=9)
        s.search_reads(chars, ends)


@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("hyper_shift", [None, "3"], ids=["hyper25", "hyper3"])
def test_wide_index_kernels(genome_index, monkeypatch, algo, hyper_shift):
    """64-bit node-index kernels (needed from 2^32 nodes) on a small index; with
    hyperblocks of 8 lines (768 nodes) the counts cross a hyperblock base on
    most steps, as those of a 4 G-node index do every 2^25 lines."""
    g, b, path, orc = genome_index
    say(f"forced 64-bit index algo={algo} hyper_shift={hyper_shift}")
    monkeypatch.setenv("SBWT_AMD_FORCE_WIDE", "1")
    if hyper_shift:
        monkeypatch.setenv("SBWT_AMD_HYPER_SHIFT", hyper_shift)
    idx = S.Index.load(path)
    assert idx.wide
    chars, ends = _reads(g, 2000, 150, seed=21)
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(algo)
    assert np.array_equal(s.search_reads(chars, ends), orc.index_reads(chars, ends))


def test_long_reads_span_many_chunks(tmp_path):
    """Reads of up to 20 kbp (many 256-k-mer work items) from similar genomes."""
    say("long reads, many chunks")
    root = W.genome_random(9, 60_000)
    gens = [root] + [W.genome_mutate(root, s, 0.02, 0.001, 10) for s in (1, 2, 3)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(tmp_path / "l.sbwt")
    b.write(path)
    orc = O.OracleSbwt(path)
    idx = S.Index.load(path)
    lens = W.lognormal_lengths(4, 300, 2000, 1.2, 31, 20_000)
    chars, rends, _ = W.reads(6, codes, ends, lens, 0.05, 0.01, 0.0005)
    chars = chars.tobytes()
    s = idx.searcher(len(chars), len(rends))
    s.set_algorithm(S.ALGO_STREAMING)
    s.set_stats(True)
    res = s.search_reads(chars, rends)
    st = s.last_stats()
    assert np.array_equal(res, orc.index_reads(chars, rends))
    assert st["streamed"] > st["full_searches"]  # most k-mers step from their predecessor
    assert S.Searcher.resolved(st) == len(res)


def test_streaming_without_suffix_group_starts(tmp_path):
    """Index built from the ACGT vectors alone: steps from nodes that do not start
    their suffix group fall back to full searches; results are unchanged. With
    the starts, such steps resolve from the group start (in the node's line or
    an earlier one) and never fall back."""
    say("streaming without suffix-group starts")
    root = W.genome_random(12, 40_000)
    gens = [root] + [W.genome_mutate(root, s, 0.03) for s in (1, 2, 3, 4)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(tmp_path / "v.sbwt")
    b.write(path)
    orc = O.OracleSbwt(path)
    chars, rends, _ = W.reads(7, codes, ends, np.full(3000, 150, np.uint32), 0.05, 0.005, 0.0)
    chars = chars.tobytes()
    exp = orc.index_reads(chars, rends)
    for sgs in (None, b.bits(4)):
        idx = S.Index.create([b.bits(c) for c in range(4)], b.n, 31, suffix_group_starts=sgs)
        s = idx.searcher(len(chars), len(rends))
        s.set_algorithm(S.ALGO_STREAMING)
        s.set_stats(True)
        assert np.array_equal(s.search_reads(chars, rends), exp)
        st = s.last_stats()
        if sgs is None:
            assert st["stream_fallbacks"] > 0
        else:
            assert st["stream_fallbacks"] == 0 and st["group_steps"] > 0
            assert st["stream_extra_lines"] <= st["group_steps"]
            # every k-mer of the genomes: includes groups that begin in the line
            # before the node's own line
            gchars = W.ACGT[codes].tobytes()
            s2 = idx.searcher(len(gchars), len(ends))
            s2.set_algorithm(S.ALGO_STREAMING)
            s2.set_stats(True)
            assert np.array_equal(s2.search_reads(gchars, ends), orc.index_reads(gchars, ends))
            assert s2.last_stats()["stream_extra_lines"] > 0


def test_create_from_bit_vectors_matches_load(genome_index):
    g, b, path, orc = genome_index
    say("create from bit vectors")
    idx = S.Index.create([b.bits(c) for c in range(4)], b.n, 31)
    assert idx.c_map == orc.c_map()
    chars, ends = _reads(g, 1000, 150, seed=8)
    s = idx.searcher(len(chars), len(ends))
    assert np.array_equal(s.search_reads(chars, ends), orc.index_reads(chars, ends))


@pytest.mark.parametrize("cap", [None, "2", "0"], ids=["table", "table-cap2", "walk-only"])
@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("d", [1, 3, 20])
def test_key_kmer_move_parity(d, algo, cap, tmp_path, monkeypatch):
    """-k: move every found k-mer to its key k-mer (IndexSearcher.cuh:60-70). The
    move reads a per-node destination table built at load; walks longer than its
    cap (lowered here through SBWT_AMD_KEY_TABLE_STEPS) are walked at query time."""
    if cap is not None:
        monkeypatch.setenv("SBWT_AMD_KEY_TABLE_STEPS", cap)
    say(f"key k-mer walk d={d} algo={algo} cap={cap}")
    _key_move_case(d, algo, tmp_path)


@pytest.mark.parametrize("cap", [None, "0"], ids=["table", "walk-only"])
@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
def test_key_kmer_move_wide_small_hyperblocks(algo, cap, tmp_path, monkeypatch):
    """The key move on 64-bit node indexes with hyperblocks of 8 lines: the mark
    and letter counts of every line add a hyperblock base."""
    monkeypatch.setenv("SBWT_AMD_FORCE_WIDE", "1")
    monkeypatch.setenv("SBWT_AMD_HYPER_SHIFT", "3")
    if cap is not None:
        monkeypatch.setenv("SBWT_AMD_KEY_TABLE_STEPS", cap)
    say(f"key k-mer walk wide hyper3 algo={algo} cap={cap}")
    _key_move_case(3, algo, tmp_path)


def _key_move_case(d, algo, tmp_path):
    root = W.genome_random(1, 30_000)
    gens = [root] + [W.genome_mutate(root, s, 0.02) for s in (2, 3, 4)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    col = b.colors(np.arange(len(gens), dtype=np.uint32), d=d)
    sp, cp = str(tmp_path / "k.sbwt"), str(tmp_path / "k.tcolors")
    b.write(sp)
    col.write(cp)
    orc = O.OracleSbwt(sp)
    km = col.key_marks()
    idx = S.Index.load(sp, cp)
    assert idx.has_key_marks
    if algo == S.ALGO_STREAMING:  # fresh searches from the k-mer table (none when wide)
        assert idx.build_kmer_table() != idx.wide
    chars, ends2, _ = W.reads(5, codes, ends, np.full(2000, 120, np.uint32), 0.2, 0.01, 0.001)
    chars = chars.tobytes()
    s = idx.searcher(len(chars), len(ends2), move_to_key_kmer=True)
    s.set_algorithm(algo)
    res = s.search_reads(chars, ends2)
    exp = orc.index_reads(chars, ends2, key_marks=np.concatenate([km, np.zeros(8, np.uint64)]))
    assert np.array_equal(res, exp)
    assert not np.any(res == O.WALK_STUCK)


def test_device_buffers_with_torch(genome_index):
    """Device-resident form: torch allocates, the library computes."""
    import torch
    g, b, path, orc = genome_index
    say("device buffers via torch")
    idx = S.Index.load(path)
    chars, ends = _reads(g, 4000, 150, seed=33)
    dc = torch.frombuffer(bytearray(chars), dtype=torch.uint8).cuda()
    de = torch.from_numpy(ends.astype(np.int64)).cuda()
    dr = torch.empty(len(chars), dtype=torch.int64, device="cuda")
    s = idx.searcher(len(chars), len(ends))
    nk = s.search_reads_device(dc.data_ptr(), len(chars), de.data_ptr(), len(ends),
                               dr.data_ptr(), want_count=True)
    assert nk == 4000 * 120
    got = dr[:nk].cpu().numpy().view(np.uint64)
    assert np.array_equal(got, orc.index_reads(chars, ends))
    t = s.last_timings()
    assert t["search_ms"] > 0


@pytest.mark.parametrize("algo", ALGOS, ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("offset_words", [1, 3, 7])
def test_output_at_any_word_offset(genome_index, algo, offset_words):
    """Results staged per 64-byte group land right for an output pointer at any
    8-byte offset, and nothing outside the k-mers' slots is written."""
    import torch
    g, b, path, orc = genome_index
    say(f"output offset {offset_words} words algo={algo}")
    idx = S.Index.load(path)
    lens = np.array([31, 45, 150, 97, 31, 300, 64, 200] * 40, np.uint32)
    chars, ends, _ = W.reads(41 + offset_words, g, np.array([len(g)], np.uint64), lens, 0.1, 0.01,
                             0.002)
    chars = chars.tobytes()
    exp = orc.index_reads(chars, ends)
    dc = torch.frombuffer(bytearray(chars), dtype=torch.uint8).cuda()
    de = torch.from_numpy(ends.astype(np.int64)).cuda()
    sentinel = -7
    buf = torch.full((len(exp) + 16,), sentinel, dtype=torch.int64, device="cuda")
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(algo)
    nk = s.search_reads_device(dc.data_ptr(), len(chars), de.data_ptr(), len(ends),
                               buf.data_ptr() + 8 * offset_words, want_count=True)
    assert nk == len(exp)
    host = buf.cpu().numpy()
    assert np.array_equal(host[offset_words:offset_words + nk].view(np.uint64), exp)
    assert (host[:offset_words] == sentinel).all()
    assert (host[offset_words + nk:] == sentinel).all()


def test_absence_checks_settle_windows(genome_index):
    """Random reads and reads with errors: one presearch lookup that finds an
    absent kp-mer settles every window holding it; results equal the oracle's and
    the census accounts for every k-mer exactly once."""
    g, b, path, orc = genome_index
    say("absence checks")
    idx = S.Index.load(path)
    assert (1 << (2 * idx.presearch_letters)) >= 8 * idx.num_nodes
    chars, ends = _reads(g, 3000, 150, seed=77, random_frac=0.3, err=0.02, n_rate=0.002)
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(S.ALGO_STREAMING)
    s.set_stats(True)
    res = s.search_reads(chars, ends)
    st = s.last_stats()
    assert np.array_equal(res, orc.index_reads(chars, ends))
    assert st["absence_hits"] > 0 and st["absent_skipped"] > st["absence_hits"]
    assert S.Searcher.resolved(st) == len(res)


@pytest.mark.parametrize("chunk,bpc", [(64, 1), (100, 3), (256, 4), (1000, 8), (65535, 2)])
def test_launch_shapes_give_identical_results(genome_index, chunk, bpc):
    """Work-item length and persistent workgroups per CU change only the speed."""
    g, b, path, orc = genome_index
    say(f"launch shape chunk={chunk} bpc={bpc}")
    idx = S.Index.load(path)
    lens = np.concatenate([np.full(500, 150, np.uint32), W.uniform_lengths(3, 100, 31, 3000)])
    chars, ends, _ = W.reads(61, g, np.array([len(g)], np.uint64), lens, 0.2, 0.01, 0.001)
    chars = chars.tobytes()
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(S.ALGO_STREAMING)
    s.set_launch(chunk, bpc)
    assert s.launch() == (chunk, bpc)
    assert np.array_equal(s.search_reads(chars, ends), orc.index_reads(chars, ends))


def test_launch_shape_bounds(genome_index):
    g, b, path, orc = genome_index
    s = S.Index.load(path).searcher(1000, 10)
    for bad in ((63, 4), (65536, 4), (256, 0), (256, 9)):
        with pytest.raises(S.SbwtError):
            s.set_launch(*bad)


def test_autotune_keeps_results(genome_index):
    import torch
    g, b, path, orc = genome_index
    say("autotune")
    idx = S.Index.load(path)
    chars, ends = _reads(g, 3000, 150, seed=91)
    dc = torch.frombuffer(bytearray(chars), dtype=torch.uint8).cuda()
    de = torch.from_numpy(ends.astype(np.int64)).cuda()
    dr = torch.empty(len(chars), dtype=torch.int64, device="cuda")
    s = idx.searcher(len(chars), len(ends))
    ch, bpc, ms = s.autotune(dc.data_ptr(), len(chars), de.data_ptr(), len(ends), dr.data_ptr())
    assert s.launch() == (ch, bpc) and ch in (128, 256, 512, 1024) and bpc in (2, 3, 4, 5, 6, 7, 8)
    assert ms > 0
    exp = orc.index_reads(chars, ends)
    assert np.array_equal(dr[:len(exp)].cpu().numpy().view(np.uint64), exp)
    assert np.array_equal(s.search_reads(chars, ends), exp)


def test_auto_algorithm_by_batch_size(genome_index):
    """The default searches per k-mer below 2^22 chars and streams above (work
    items shrunk when the batch is small for the chip); streaming-adaptive
    streams from 2^16 chars. The results are the same."""
    g, b, path, orc = genome_index
    say("auto algorithm")
    idx = S.Index.load(path)
    tiny_c, tiny_e = _reads(g, 300, 150, seed=70)
    small_c, small_e = _reads(g, 1000, 150, seed=71)
    big_c, big_e = _reads(g, 30_000, 150, seed=72)
    assert len(tiny_c) < (1 << 16) <= len(small_c) < (1 << 22) <= len(big_c)
    s = idx.searcher(len(big_c), len(big_e))
    s.set_stats(True)
    expect = {0: orc.index_reads(tiny_c, tiny_e), 1: orc.index_reads(small_c, small_e),
              2: orc.index_reads(big_c, big_e)}
    batches = {0: (tiny_c, tiny_e), 1: (small_c, small_e), 2: (big_c, big_e)}
    for algo, streams in ((S.ALGO_AUTO, (False, False, True)),
                          (S.ALGO_STREAMING_ADAPTIVE, (False, True, True))):
        s.set_algorithm(algo)
        for i in range(3):
            c, e = batches[i]
            assert np.array_equal(s.search_reads(c, e), expect[i]), (algo, i)
            assert (s.last_stats()["streamed"] > 0) == streams[i], (algo, i)


@pytest.mark.parametrize("algo,launch", [(S.ALGO_STREAMING, (64, 4)),
                                         (S.ALGO_STREAMING_ADAPTIVE, None)],
                         ids=["items-64", "adaptive-items"])
def test_key_kmer_move_across_work_items(algo, launch, tmp_path):
    """The key pass resolves chains of same-key k-mers per work item; reads of
    up to 3 kbp split into many items (64 k-mers, or the adaptive 8+ of a
    small batch) cut chains at every item start, where the search restarts."""
    say(f"key move across work items algo={algo} launch={launch}")
    root = W.genome_random(41, 40_000)
    gens = [root] + [W.genome_mutate(root, s, 0.01) for s in (42, 43)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    col = b.colors(np.arange(len(gens), dtype=np.uint32), d=7)
    sp, cp = str(tmp_path / "w.sbwt"), str(tmp_path / "w.tcolors")
    b.write(sp)
    col.write(cp)
    orc = O.OracleSbwt(sp)
    km = np.concatenate([col.key_marks(), np.zeros(8, np.uint64)])
    idx = S.Index.load(sp, cp)
    lens = W.uniform_lengths(44, 300, 31, 3000)
    chars, rends, _ = W.reads(45, codes, ends, lens, 0.1, 0.003, 0.0)
    chars = chars.tobytes()
    s = idx.searcher(len(chars), len(rends), move_to_key_kmer=True)
    s.set_algorithm(algo)
    if launch:
        s.set_launch(*launch)
    res = s.search_reads(chars, rends)
    assert np.array_equal(res, orc.index_reads(chars, rends, key_marks=km))


@pytest.mark.parametrize("table", ["1", "0", "sliced"], ids=["kmer-table", "no-table",
                                                              "sliced-build"])
@pytest.mark.parametrize("algo", ALGOS + [S.ALGO_STREAMING_ADAPTIVE],
                         ids=["streaming", "per-kmer", "adaptive"])
def test_kmer_table_and_full_search_agree(genome_index, algo, table, monkeypatch):
    """Fresh searches start from the k-mer table (every whole-k-mer node by its
    code) when the index has one, else from the presearch interval and SBWT
    steps; both give the oracle's results, with reads that miss often. The
    sliced build (level lists of 4096 entries against ~0.4 M nodes: the first
    slice overflows and is split into quarters down to slices that fit) gives
    the same table."""
    g, b, path, orc = genome_index
    say(f"k-mer table={table} algo={algo}")
    if table == "sliced":
        monkeypatch.setenv("SBWT_AMD_KMER_TABLE_LIST_CAP", "4096")
    idx = S.Index.load(path)
    if table != "0":
        assert idx.build_kmer_table()
    chars, ends = _reads(g, 3000, W.uniform_lengths(52, 3000, 20, 300), seed=51, random_frac=0.2,
                         err=0.02, n_rate=0.005)
    s = idx.searcher(len(chars), len(ends))
    s.set_algorithm(algo)
    s.set_stats(True)
    res = s.search_reads(chars, ends)
    assert np.array_equal(res, orc.index_reads(chars, ends))
    st = s.last_stats()
    assert S.Searcher.resolved(st) == len(res) or algo == S.ALGO_PER_KMER
