"""Color search at scale on the MI355X: the batched form's work queue of
(read, k-mer range) items, and color counts beyond the LDS (20,000 and 70,000
colors, counted in per-wave global counters), against the oracle.

The reference sizes its color results as reads x num_colors with no cap
(src/ColorSearcher/ColorSearcher.cpp:25-33,51); so does this library.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W
from colorlib import K, ColoredIndex, colored_seq_ids, kmer_ends_of, read_stats

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def expected_lists(ci, res, ke, tau, inc_nf=False, inc_inv=False, first_read_quirk=False):
    """Oracle counts thresholded per read (quirk 2 included), vectorized."""
    n = len(ke)
    nc = ci.num_colors
    counts = np.zeros(max(n, 1) * nc, np.uint64)
    O.lib().orc_color_counts(C.byref(ci.oc), O.p64(np.ascontiguousarray(res, np.uint64)),
                             O.p64(np.ascontiguousarray(ke, np.uint64)), n, O.p64(counts), 8)
    counts = counts[:n * nc].reshape(n, nc)
    stats = read_stats(res, ke)
    slots = colored_seq_ids(stats)
    colored = [r for r, st in enumerate(stats) if st[0] > 0]
    out = []
    for r, (f, nf, inv) in enumerate(stats):
        if first_read_quirk and r == 0:
            f, nf, inv = [(x - 1) % (1 << 64) for x in (f, nf, inv)]
        mn = O.lib().orc_color_min_hits(f, nf, inv, tau, int(inc_nf), int(inc_inv))
        q = slots[r]
        out.append([] if mn == 0 or q >= len(colored) else
                   np.flatnonzero(counts[colored[q]] >= mn).tolist())
    return out, counts


@pytest.fixture(scope="module")
def long_reads_index(tmp_path_factory):
    ci = ColoredIndex(tmp_path_factory.mktemp("items"), n_genomes=10, length=60_000, d=12,
                      sub=0.02)
    lens = np.concatenate([W.lognormal_lengths(31, 120, 3000, 1.2, 31, 40_000),
                           np.full(300, 150, np.uint32), np.array([0, 10, 31, 32], np.uint32)])
    lens = lens[W.permutation(32, len(lens)).astype(np.int64)]
    chars, ends = ci.reads(len(lens), lens, seed=33, random_frac=0.1, err=0.01)
    res = ci.index_results(chars, ends)
    return ci, res, kmer_ends_of(ends)


@pytest.mark.parametrize("items", [64, 1024, 4096])
@pytest.mark.parametrize("tau,inc_nf,inc_inv,quirk", [(0.7, False, False, True),
                                                      (0.5, True, True, False)])
def test_work_items_equal_per_read_and_oracle(long_reads_index, items, tau, inc_nf, inc_inv,
                                              quirk):
    """Reads up to 40 kbp split into items of 64 to 4096 k-mers: the partial counts
    of a long read meet in its slot and the last item thresholds them."""
    ci, res, ke = long_reads_index
    say(f"work items {items} tau={tau} nf={inc_nf} inv={inc_inv} quirk={quirk}")
    col = S.ColorIndex(ci.colors_path)
    p = S.color_params(tau, inc_nf, inc_inv, quirk)
    exp, _ = expected_lists(ci, res, ke, tau, inc_nf, inc_inv, quirk)
    per_read = col.searcher(len(res) + 64, len(ke))
    lists0, st0, _, _ = per_read.pseudoalign(res, ke, p)
    assert lists0 == exp
    cs = col.searcher(len(res) + 64, len(ke))
    cs.set_work_items(items)
    for rep in range(2):  # the long-read slots are clean again after a batch
        lists, st, off, ids = cs.pseudoalign(res, ke, p)
        assert lists == exp, rep
        assert np.array_equal(st, st0)
    assert max(int(ke[i]) - (int(ke[i - 1]) if i else 0) for i in range(len(ke))) > 8 * items


def test_autotuned_work_items_equal_oracle(long_reads_index):
    """The work-item autotuner times every shape on the batch, keeps one of them, and
    both its own runs and the searcher's later batches give the oracle's lists."""
    import torch
    ci, res, ke = long_reads_index
    say("work-item autotune")
    col = S.ColorIndex(ci.colors_path)
    p = S.color_params(0.7)
    exp, _ = expected_lists(ci, res, ke, 0.7)
    cs = col.searcher(len(res) + 64, len(ke))
    dr = torch.from_numpy(res.view(np.int64)).cuda()
    dk = torch.from_numpy(ke.view(np.int64)).cuda()
    d_off = torch.empty(len(ke) + 1, dtype=torch.int64, device="cuda")
    cap = len(ke) * ci.num_colors
    d_ids = torch.empty(cap, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    items, ms = cs.autotune(dr.data_ptr(), dk.data_ptr(), len(ke), p, d_off.data_ptr(),
                            d_ids.data_ptr(), cap)
    say(f"  kept {items} k-mers per item, {ms:.3f} ms")
    assert items in (0, 256, 512, 1024, 2048, 4096) and ms > 0
    off = d_off.cpu().numpy().view(np.uint64)
    ids = d_ids.cpu().numpy().view(np.uint32)
    assert [ids[int(off[r]):int(off[r + 1])].tolist() for r in range(len(ke))] == exp
    lists, _, _, _ = cs.pseudoalign(res, ke, p)
    assert lists == exp


def _many_colors_index(tmp, n_genomes, length, n_colors, seed):
    # genome g gets color g * stride: the ids spread over [0, n_colors)
    stride = (n_colors - 1) // (n_genomes - 1)
    colors = np.arange(n_genomes, dtype=np.uint32) * np.uint32(stride)
    colors[-1] = n_colors - 1
    return ColoredIndex(tmp, n_genomes=n_genomes, length=length, sub=0.004, d=10, seed=seed,
                        colors=colors)


@pytest.mark.parametrize("n_colors,n_genomes,length", [(20_000, 2000, 2000),
                                                        (70_000, 5000, 1000)],
                         ids=["20k-colors", "70k-colors"])
def test_many_colors_equal_oracle(tmp_path, n_colors, n_genomes, length):
    say(f"{n_colors} colors over {n_genomes} genomes x {length} bp")
    ci = _many_colors_index(tmp_path, n_genomes, length, n_colors, seed=7)
    assert ci.num_colors == n_colors
    assert ci.colors.num_dense > 0 and ci.colors.num_sets > ci.colors.num_dense
    chars, ends = ci.reads(300, 120, seed=8, random_frac=0.1, err=0.005)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    col = S.ColorIndex(ci.colors_path)
    assert col.num_colors == n_colors
    for items in (0, 128):
        for tau, inc_nf in ((0.7, False), (0.3, True)):
            say(f"{n_colors} colors items={items} tau={tau} nf={inc_nf}")
            exp, counts = expected_lists(ci, res, ke, tau, inc_nf)
            cs = col.searcher(len(res) + 64, len(ke))
            cs.set_work_items(items)
            lists, _, _, _ = cs.pseudoalign(res, ke, S.color_params(tau, inc_nf))
            assert lists == exp
            assert sum(len(x) for x in lists) > 0
    # raw per-read counts through the device counts interface
    import torch
    say(f"{n_colors} colors: raw counts")
    d_res = torch.from_numpy(res.view(np.int64)).cuda()
    d_ke = torch.from_numpy(ke.view(np.int64)).cuda()
    d_cnt = torch.zeros(len(ke) * n_colors, dtype=torch.int32, device="cuda")
    cs = col.searcher(len(res) + 64, len(ke))
    cs.counts_device(d_res.data_ptr(), d_ke.data_ptr(), len(ke), d_cnt.data_ptr())
    torch.cuda.synchronize()
    got = d_cnt.cpu().numpy().view(np.uint32).reshape(len(ke), n_colors)
    assert np.array_equal(got.astype(np.uint64), counts)
    # the reference's warp-padded interface (u64 counts per colored read)
    say(f"{n_colors} colors: warped interface")
    n = len(ke)
    warped = np.zeros(len(res) + 64 * n + 64, np.uint64)
    wi = np.zeros(n + 2, np.uint64)
    ni = np.zeros(1, np.uint64)
    csid = np.zeros(n + 1, np.uint64)
    f, nf, inv = (np.zeros(n + 1, np.uint64) for _ in range(3))
    nw = O.lib().orc_warp_layout(O.p64(res), O.p64(ke), n, 64, O.p64(warped), O.p64(wi),
                                 O.p64(ni), O.p64(csid), O.p64(f), O.p64(nf), O.p64(inv))
    n_int = int(ni[0])
    cw = col.searcher(int(nw) + 64, n_int)
    wcounts = cw.search(warped[:nw], wi[:n_int])
    colored = [r for r in range(n) if int(f[r]) > 0]
    assert np.array_equal(wcounts.reshape(n_int - 1, n_colors), counts[colored])


@pytest.mark.parametrize("algo", [S.ALGO_STREAMING, S.ALGO_PER_KMER], ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("n_genomes", [8, 80], ids=["masks-8-colors", "ranks-80-colors"])
def test_fused_color_table_equals_oracle(tmp_path, n_genomes, algo):
    """The fused path: the index's color table gives each k-mer its key's color mask
    (at most 61 colors) or key rank; the index stage writes those entries and the
    color stage counts them. Entries and color lists equal the oracle's."""
    say(f"fused color table: {n_genomes} colors algo={algo}")
    ci = ColoredIndex(tmp_path, n_genomes=n_genomes, length=12_000, sub=0.02, d=9, seed=5)
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    col = S.ColorIndex(ci.colors_path)
    kind = idx.build_color_table(col)
    assert kind == (1 if n_genomes <= 61 else 2)
    lens = np.concatenate([np.full(600, 120, np.uint32), W.uniform_lengths(61, 60, 31, 6000)])
    chars, ends = ci.reads(len(lens), lens, seed=62, random_frac=0.1, err=0.01)
    keys = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    # expected table entries from the oracle's key nodes
    km = ci.key_marks
    bits = np.unpackbits(km.view(np.uint8), bitorder="little").astype(np.int64)
    ranks = np.concatenate([[0], np.cumsum(bits)])
    exp_entries = keys.copy()
    found = keys < np.uint64(ci.sbwt.n)
    if kind == 2:
        exp_entries[found] = ranks[keys[found].astype(np.int64)].astype(np.uint64)
    else:
        for i in np.flatnonzero(found):
            m = 0
            for c in ci.color_set(int(keys[i])):
                m |= 1 << c
            exp_entries[i] = np.uint64(m)
    s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
    s.set_algorithm(algo)
    s.set_color_output(True)
    got = s.search_reads(chars, ends)
    assert np.array_equal(got, exp_entries)
    s.set_color_output(False)
    assert np.array_equal(s.search_reads(chars, ends), keys)
    for items in (0, 64):
        for tau, inc_nf, quirk in ((0.7, False, True), (0.4, True, False)):
            exp, _ = expected_lists(ci, keys, ke, tau, inc_nf, False, quirk)
            cs = col.searcher(len(got) + 64, len(ke))
            cs.set_work_items(items)
            cs.set_input(idx)
            lists, _, _, _ = cs.pseudoalign(got, ke, S.color_params(tau, inc_nf, False, quirk))
            assert lists == exp, (items, tau)
    # a color searcher of another color index refuses this index's entries
    other = S.ColorIndex(ci.colors_path)
    with pytest.raises(S.SbwtError):
        other.searcher(100, 10).set_input(idx)


@pytest.mark.parametrize("n_genomes", [8, 80], ids=["masks-8-colors", "ranks-80-colors"])
def test_fused_entries_with_other_index_marks_match_per_kmer(tmp_path, n_genomes):
    """The streaming key pass takes a marked node's key rank from the search (its line's
    mark rank: the color-table entry with ranks, the index of its mask with masks) only
    when the index's own key marks are the color index's: with one more node marked in
    the index, those ranks differ, and the streaming fused entries must still equal the
    per-k-mer fused search's, which reads the table."""
    say(f"fused entries, {n_genomes} colors: index marks differ from the color index's")
    ci = ColoredIndex(tmp_path, n_genomes=n_genomes, length=12_000, sub=0.02, d=9, seed=5)
    km = ci.key_marks.copy()
    bits = np.unpackbits(km.view(np.uint8), bitorder="little")[:ci.sbwt.n]
    extra = int(np.flatnonzero(bits == 0)[0])
    km.view(np.uint8)[extra // 8] |= np.uint8(1 << (extra % 8))
    idx = S.Index.create([ci.sbwt.bits(c) for c in range(4)], ci.sbwt.n, K, key_marks=km,
                         suffix_group_starts=ci.sbwt.bits(4))
    col = S.ColorIndex(ci.colors_path)
    assert idx.build_color_table(col) == (1 if n_genomes <= 61 else 2)
    lens = np.concatenate([np.full(600, 120, np.uint32), W.uniform_lengths(63, 60, 31, 6000)])
    chars, ends = ci.reads(len(lens), lens, seed=64, random_frac=0.1, err=0.01)
    got = {}
    for algo in (S.ALGO_STREAMING, S.ALGO_PER_KMER):
        s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
        s.set_algorithm(algo)
        s.set_color_output(True)
        got[algo] = s.search_reads(chars, ends)
    assert np.array_equal(got[S.ALGO_STREAMING], got[S.ALGO_PER_KMER])


@pytest.mark.parametrize("algo", [S.ALGO_STREAMING, S.ALGO_PER_KMER], ids=["streaming", "per-kmer"])
@pytest.mark.parametrize("n_genomes", [8, 80], ids=["masks-8-colors", "ranks-80-colors"])
def test_fused_key_chains_equal_oracle(tmp_path, n_genomes, algo):
    """Key-chain output (the search leaves the key pass to the color stage, which
    resolves each chain while it counts, right to left; work items of 64 k-mers cut
    chains at their ends) gives the oracle's lists, per read and per work item."""
    say(f"fused key chains: {n_genomes} colors algo={algo}")
    ci = ColoredIndex(tmp_path, n_genomes=n_genomes, length=12_000, sub=0.02, d=9, seed=5)
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    col = S.ColorIndex(ci.colors_path)
    idx.build_color_table(col)
    lens = np.concatenate([np.full(600, 120, np.uint32), W.uniform_lengths(61, 60, 31, 6000)])
    chars, ends = ci.reads(len(lens), lens, seed=62, random_frac=0.1, err=0.01)
    keys = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
    s.set_algorithm(algo)
    s.set_color_output(S.COLOR_OUTPUT_CHAINS)
    raw = s.search_reads(chars, ends)
    for items in (0, 64):
        for tau, inc_nf, quirk in ((0.7, False, True), (0.4, True, False)):
            say(f"  items={items} tau={tau} nf={inc_nf}")
            exp, _ = expected_lists(ci, keys, ke, tau, inc_nf, False, quirk)
            cs = col.searcher(len(raw) + 64, len(ke))
            cs.set_work_items(items)
            cs.set_input(idx, chains=True)
            lists, _, _, _ = cs.pseudoalign(raw, ke, S.color_params(tau, inc_nf, False, quirk))
            assert lists == exp, (items, tau)


def test_fused_key_chains_with_other_index_marks(tmp_path):
    """Key chains over an index whose marks differ from the color index's (no ranks
    from the search: every chain end reads the color table) give the same lists as
    the color-table entries."""
    say("fused key chains: index marks differ from the color index's")
    ci = ColoredIndex(tmp_path, n_genomes=80, length=12_000, sub=0.02, d=9, seed=5)
    km = ci.key_marks.copy()
    bits = np.unpackbits(km.view(np.uint8), bitorder="little")[:ci.sbwt.n]
    extra = int(np.flatnonzero(bits == 0)[0])
    km.view(np.uint8)[extra // 8] |= np.uint8(1 << (extra % 8))
    idx = S.Index.create([ci.sbwt.bits(c) for c in range(4)], ci.sbwt.n, K, key_marks=km,
                         suffix_group_starts=ci.sbwt.bits(4))
    col = S.ColorIndex(ci.colors_path)
    idx.build_color_table(col)
    lens = np.concatenate([np.full(600, 120, np.uint32), W.uniform_lengths(63, 60, 31, 6000)])
    chars, ends = ci.reads(len(lens), lens, seed=64, random_frac=0.1, err=0.01)
    ke = kmer_ends_of(ends)
    got = {}
    for mode in (S.COLOR_OUTPUT_ENTRIES, S.COLOR_OUTPUT_CHAINS):
        s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
        s.set_algorithm(S.ALGO_STREAMING)
        s.set_color_output(mode)
        res = s.search_reads(chars, ends)
        cs = col.searcher(len(res) + 64, len(ke))
        cs.set_input(idx, chains=mode == S.COLOR_OUTPUT_CHAINS)
        got[mode], _, _, _ = cs.pseudoalign(res, ke, S.color_params(0.7, False, False, True))
    assert got[S.COLOR_OUTPUT_ENTRIES] == got[S.COLOR_OUTPUT_CHAINS]
