"""The benchmark workloads C1-C5 (sbwt-search_amd/bindings/configs.py) on the
MI355X, checked against the oracle.

C1 runs at full size against the full oracle output. C2 runs at full size
(95 M-node index, 10 M reads): the oracle checks a strided sample of 10^5
reads, and at full size the streaming and per-k-mer kernels (independent
algorithms) must agree on all 1.2 G k-mers. C3, C4 and C5 run their recipes
at sizes the oracle covers in seconds: C4's batched form and its plain form
must give identical full outputs, equal to the oracle's; C5's served results
must equal each request's result alone and the oracle's.
"""
import os

import numpy as np
import pytest

import configs
import oracle_py as O
import sbwt_amd as S
import workloads as W
from colorlib import colored_seq_ids, kmer_ends_of, read_stats

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def _threads():
    return min(16, os.cpu_count() or 1)


def oracle_lists(oc_counts, results, kmer_ends, tau, inc_nf=False, inc_inv=False,
                 first_read_quirk=False):
    """Vectorized form of colorlib.threshold_lists (quirk 2 included): per read the
    sorted colors of its colored slot's counts that reach ceil(seq_size * tau)."""
    stats = read_stats(results, kmer_ends)
    slots = colored_seq_ids(stats)
    colored = [r for r, st in enumerate(stats) if st[0] > 0]
    out = []
    for r, (f, nf, inv) in enumerate(stats):
        if first_read_quirk and r == 0:
            f, nf, inv = [(x - 1) % (1 << 64) for x in (f, nf, inv)]
        mn = O.lib().orc_color_min_hits(f, nf, inv, tau, int(inc_nf), int(inc_inv))
        q = slots[r]
        if mn == 0 or q >= len(colored):
            out.append([])
            continue
        out.append(np.flatnonzero(oc_counts[colored[q]] >= mn).tolist())
    return out


class Colored:
    """A workload's SBWT + color index (one color per genome) on disk, loaded by the
    oracle and the library."""

    def __init__(self, tmp, codes, ends, d=20):
        import ctypes as C
        self.sbwt = W.BuiltSbwt(codes, ends, configs.K)
        self.col = self.sbwt.colors(np.arange(len(ends), dtype=np.uint32), d=d)
        self.sbwt_path = str(tmp / "w.sbwt")
        self.colors_path = str(tmp / "w.tcolors")
        self.sbwt.write(self.sbwt_path)
        self.col.write(self.colors_path)
        self.key_marks = np.concatenate([self.col.key_marks(), np.zeros(8, np.uint64)])
        self.orc = O.OracleSbwt(self.sbwt_path)
        self.oc = O.Colors()
        if O.lib().orc_colors_load(self.colors_path.encode(), C.byref(self.oc)):
            raise RuntimeError("oracle could not load the color index")
        self.num_colors = int(self.oc.num_colors)

    def index_results(self, chars, ends):
        return self.orc.index_reads(chars, ends, key_marks=self.key_marks, threads=_threads())

    def counts(self, results, kmer_ends):
        n = len(kmer_ends)
        out = np.zeros(max(n, 1) * self.num_colors, np.uint64)
        O.lib().orc_color_counts(O.C.byref(self.oc), O.p64(np.ascontiguousarray(results, np.uint64)),
                                 O.p64(np.ascontiguousarray(kmer_ends, np.uint64)), n, O.p64(out),
                                 _threads())
        return out[:n * self.num_colors].reshape(n, self.num_colors)


# ---------------------------------------------------------------- C1

@pytest.fixture(scope="module")
def c1_case(tmp_path_factory):
    wl = configs.c1()
    b = W.BuiltSbwt(wl.genomes, wl.genome_ends, configs.K)
    path = str(tmp_path_factory.mktemp("c1") / "c1.sbwt")
    b.write(path)
    orc = O.OracleSbwt(path)
    chars = wl.chars.tobytes()
    exp = orc.index_reads(chars, wl.read_ends, threads=_threads())
    return wl, b, path, exp


@pytest.mark.parametrize("table", [False, True], ids=["no-table", "kmer-table"])
@pytest.mark.parametrize("algo", [S.ALGO_STREAMING, S.ALGO_PER_KMER, S.ALGO_AUTO,
                                  S.ALGO_STREAMING_ADAPTIVE],
                         ids=["streaming", "per-kmer", "auto", "adaptive"])
def test_c1_full_size_equals_oracle(c1_case, algo, table):
    wl, b, path, exp = c1_case
    say(f"C1 full size algo={algo} table={table}")
    assert wl.n_reads == 10_000 and len(exp) == 10_000 * 120
    idx = S.Index.load(path)
    if table:
        assert idx.build_kmer_table()
    s = idx.searcher(wl.n_chars, wl.n_reads)
    s.set_algorithm(algo)
    res = s.search_reads(wl.chars.tobytes(), wl.read_ends)
    assert np.array_equal(res, exp)
    # the recipe's classes are all present: found, not found and invalid (N)
    assert np.count_nonzero(res < O.INVALID) > 0.3 * len(res)
    assert np.count_nonzero(res == O.INVALID) > 0
    assert np.count_nonzero(res == O.NOT_FOUND) > 0.3 * len(res)


# ---------------------------------------------------------------- C2

def test_c2_full_size_sampled_oracle_and_kernels_agree():
    import torch
    say("C2 full size: workload and SBWT")
    wl = configs.c2()
    assert wl.n_reads == 10_000_000
    b = W.BuiltSbwt(wl.genomes, wl.genome_ends, configs.K)
    bits = [b.bits(c) for c in range(4)]
    idx = S.Index.create(bits, b.n, configs.K, suffix_group_starts=b.bits(4))
    assert idx.has_suffix_groups and not idx.suffix_groups_rejected
    assert idx.build_kmer_table()
    n_kmers = wl.n_kmers()
    assert n_kmers == 10_000_000 * 120
    dev = torch.device("cuda:0")
    d_chars = torch.from_numpy(wl.chars).to(dev)
    d_ends = torch.from_numpy(wl.read_ends.view(np.int64)).to(dev)
    out = {}
    for algo, name in ((S.ALGO_STREAMING, "streaming"), (S.ALGO_PER_KMER, "per-kmer")):
        say(f"C2 full size search algo={name}")
        s = idx.searcher(wl.n_chars, wl.n_reads)
        s.set_algorithm(algo)
        d_res = torch.empty(wl.n_chars, dtype=torch.int64, device=dev)
        nk = s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                                   d_res.data_ptr(), want_count=True)
        assert nk == n_kmers
        out[name] = d_res[:nk]
        del s
    say("C2 full size: streaming == per-k-mer on all k-mers")
    assert torch.equal(out["streaming"], out["per-kmer"])
    del out["per-kmer"]
    # strided sample of 10^5 reads (every 100th) against the oracle
    say("C2 sampled oracle check")
    sample = np.arange(0, wl.n_reads, 100)
    orc = O.OracleSbwt(acgt=bits, num_bits=b.n, k=configs.K)
    ends = wl.read_ends.astype(np.int64)
    starts = np.concatenate([[0], ends[:-1]])
    kstarts = np.concatenate([[0], np.cumsum(np.maximum(ends - starts - configs.K + 1, 0))[:-1]])
    pieces, pends, kidx = [], [], []
    tot = 0
    for r in sample:
        pieces.append(wl.chars[starts[r]:ends[r]])
        tot += ends[r] - starts[r]
        pends.append(tot)
        nk = max(0, ends[r] - starts[r] - configs.K + 1)
        kidx.append(np.arange(kstarts[r], kstarts[r] + nk))
    chars = np.concatenate(pieces).tobytes()
    exp = orc.index_reads(chars, np.array(pends, np.uint64), threads=_threads())
    kidx = torch.from_numpy(np.concatenate(kidx)).to(dev)
    got = out["streaming"].index_select(0, kidx).cpu().numpy().view(np.uint64)
    assert len(exp) == len(sample) * 120
    assert np.array_equal(got, exp)
    found = np.count_nonzero(exp < O.INVALID) / len(exp)
    assert 0.6 < found < 0.85


# ---------------------------------------------------------------- C3

def test_c3_recipe_colors_equal_oracle(tmp_path):
    """C3's recipe (genomes from seeded roots with 1-5% substitutions, one color
    each, d = 20, reads U[50, 300] with 1% error) at 60 colors."""
    say("C3 recipe: 60 colors, 40,000 reads")
    wl = configs.c3(n_reads=40_000, n_genomes=60, n_roots=4, min_len=20_000, max_len=50_000)
    ci = Colored(tmp_path, wl.genomes, wl.genome_ends, d=20)
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    idx.build_kmer_table()
    col = S.ColorIndex(ci.colors_path)
    assert col.num_colors == 60
    chars = wl.chars.tobytes()
    exp_res = ci.index_results(chars, wl.read_ends)
    ke = kmer_ends_of(wl.read_ends)
    counts = ci.counts(exp_res, ke)
    for algo in (S.ALGO_STREAMING, S.ALGO_AUTO):
        s = idx.searcher(len(chars), wl.n_reads, move_to_key_kmer=True)
        s.set_algorithm(algo)
        res = s.search_reads(chars, wl.read_ends)
        assert np.array_equal(res, exp_res), algo
    cs = col.searcher(len(exp_res) + 64, wl.n_reads)
    for tau, inc_nf, inc_inv in ((0.7, False, False), (0.7, True, False), (0.5, True, True)):
        say(f"C3 recipe pseudoalign tau={tau} nf={inc_nf} inv={inc_inv}")
        lists, st, _, _ = cs.pseudoalign(exp_res, ke, S.color_params(tau, inc_nf, inc_inv))
        assert lists == oracle_lists(counts, exp_res, ke, tau, inc_nf, inc_inv)
    assert sum(len(x) for x in lists) > 0


# ---------------------------------------------------------------- C4

def test_c4_shape_batched_equals_plain_and_oracle(tmp_path):
    """C4's shape: log-normal reads (median 2 kbp, sigma 1.2, up to 100 kbp)
    interleaved with 150 bp reads, over C2-recipe genomes with their color index,
    through the batched and the plain form (bench.py --mode batched) on several
    batches: full outputs identical, and equal to the oracle's."""
    import torch
    import modes
    say("C4 shape: 12 genomes x 200 kbp, 3,000 log-normal + 12,000 short reads")
    wl = configs.c4(n_long=3000, n_short=12_000, n_genomes=12, length=200_000)
    lens = np.diff(np.concatenate([[0], wl.read_ends.astype(np.int64)]))
    assert lens.max() > 20_000 and np.count_nonzero(lens == 150) >= 12_000
    ci = Colored(tmp_path, wl.genomes, wl.genome_ends, d=20)
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    assert idx.build_kmer_table()
    col = S.ColorIndex(ci.colors_path)
    assert idx.build_color_table(col) == 1  # 12 colors: masks
    batch_chars = wl.n_chars // 5 + 1
    r = modes.run_batched(S, torch, wl, idx, col, batch_chars, tau=0.7, log=say, collect=True)
    assert r["batches"] >= 5
    outs = r["outputs"]
    chars = wl.chars.tobytes()
    ends = wl.read_ends.astype(np.int64)
    for b, (r0, r1) in enumerate(outs["batches"]):
        say(f"C4 shape: batch {b} against the oracle")
        c0 = int(ends[r0 - 1]) if r0 else 0
        bends = (ends[r0:r1] - c0).astype(np.uint64)
        res = ci.index_results(chars[c0:int(ends[r1 - 1])], bends)
        ke = kmer_ends_of(bends)
        exp = oracle_lists(ci.counts(res, ke), res, ke, 0.7)
        off, ids = outs["plain"][b]
        got = [ids[int(off[i]):int(off[i + 1])].tolist() for i in range(r1 - r0)]
        assert got == exp, b
        for form in ("batched", "batched_unfused"):
            assert np.array_equal(outs[form][b][0], off)
            assert np.array_equal(outs[form][b][1], ids)


# ---------------------------------------------------------------- C5

def test_c5_recipe_served_equals_alone_and_oracle(tmp_path):
    """C5's recipe (genomes from seeded roots at 3% substitutions, one color each;
    requests of N log-uniform reads of U[75, 250] bp) at 20 genomes x 100 kbp,
    through the serving loop with requests in flight together."""
    say("C5 recipe: 20 genomes x 100 kbp, 30 requests")
    codes, gends = configs.c5_genomes(n_genomes=20, n_roots=2, length=100_000)
    ci = Colored(tmp_path, codes, gends, d=20)
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    idx.build_kmer_table()
    col = S.ColorIndex(ci.colors_path)
    rng = np.random.default_rng(5)
    sizes = np.maximum(np.exp(rng.uniform(0, np.log(5000), 30)).astype(np.int64), 1)
    reqs = []
    for i, n in enumerate(sizes):
        lens = W.uniform_lengths(500 + i, int(n), 75, 250)
        c, e, _ = W.reads(500 + i, codes, gends, lens, random_frac=0.0, err_rate=0.01)
        reqs.append((c.tobytes(), e))
    params = S.color_params(0.7)
    srv = S.Server(idx, col, slots=4, max_chars_per_slot=1 << 20, max_reads_per_slot=8192,
                   move_to_key_kmer=True, params=params)
    ids = [srv.submit(c, e) for c, e in reqs]
    got = [srv.wait(i) for i in ids]
    assert srv.peak_in_flight() >= 2
    for i, ((chars, ends), g) in enumerate(zip(reqs, got)):
        res = ci.index_results(chars, ends)
        assert np.array_equal(g["kmers"], res), i
        ke = kmer_ends_of(ends)
        assert g["color_lists"] == oracle_lists(ci.counts(res, ke), res, ke, 0.7), i
    srv.close()
