"""Color search on the MI355X through the C ABI, bit-exact against the oracle."""
import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module", params=[(8, 5), (40, 20)], ids=["8colors-d5", "40colors-d20"])
def ci(request, tmp_path_factory):
    n, d = request.param
    return ColoredIndex(tmp_path_factory.mktemp("gcol"), n_genomes=n, length=20_000, d=d,
                        sub=0.03)


@pytest.fixture(scope="module")
def gpu_index(ci):
    return S.Index.load(ci.sbwt_path, ci.colors_path), S.ColorIndex(ci.colors_path)


def test_color_index_info(ci, gpu_index):
    say("color index info")
    _, col = gpu_index
    assert col.num_colors == ci.num_colors
    assert col.num_color_sets == ci.colors.num_sets
    assert col.num_nodes == ci.sbwt.n


def test_key_walk_index_stage_matches_oracle(ci, gpu_index):
    say("index stage with key walk")
    idx, _ = gpu_index
    chars, ends = ci.reads(1500, 120, seed=1)
    s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
    assert np.array_equal(s.search_reads(chars, ends), ci.index_results(chars, ends))


@pytest.mark.parametrize("tau,inc_nf,inc_inv,quirk", [(0.7, False, False, False),
                                                      (0.7, False, False, True),
                                                      (0.3, True, False, False),
                                                      (1.0, True, True, False),
                                                      (0.0, False, False, False)])
def test_pseudoalign_matches_oracle(ci, gpu_index, tau, inc_nf, inc_inv, quirk):
    say(f"pseudoalign tau={tau} nf={inc_nf} inv={inc_inv} quirk={quirk}")
    _, col = gpu_index
    chars, ends = ci.reads(800, [50 + (i * 37) % 250 for i in range(800)], seed=2)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    counts = ci.oracle_counts(res, ke)
    stats = read_stats(res, ke)
    exp = threshold_lists(counts, stats, tau, inc_nf, inc_inv, quirk)
    cs = col.searcher(len(res) + 64, len(ke))
    lists, st, offsets, ids = cs.pseudoalign(res, ke, S.color_params(tau, inc_nf, inc_inv, quirk))
    assert [tuple(x) for x in st.tolist()] == stats
    assert lists == exp
    assert sum(len(x) for x in exp) > 0 or tau == 0.0


def test_counts_device_matches_oracle(ci, gpu_index):
    import torch
    say("raw counts on device")
    _, col = gpu_index
    chars, ends = ci.reads(600, 150, seed=3)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    counts = ci.oracle_counts(res, ke)
    d_res = torch.from_numpy(res.view(np.int64)).cuda()
    d_ke = torch.from_numpy(ke.view(np.int64)).cuda()
    d_counts = torch.zeros(len(ke) * ci.num_colors, dtype=torch.int32, device="cuda")
    cs = col.searcher(len(res), len(ke))
    cs.counts_device(d_res.data_ptr(), d_ke.data_ptr(), len(ke), d_counts.data_ptr())
    torch.cuda.synchronize()
    got = d_counts.cpu().numpy().view(np.uint32).reshape(len(ke), ci.num_colors)
    assert np.array_equal(got.astype(np.uint64), counts)


def test_reference_warped_interface(ci, gpu_index):
    """ColorSearcher::search(warped idxs, warp intervals) -> u64 counts per colored read."""
    say("warped interface")
    _, col = gpu_index
    chars, ends = ci.reads(400, 200, seed=4)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    n = len(ke)
    warped = np.zeros(len(res) + 64 * n + 64, np.uint64)
    wi = np.zeros(n + 2, np.uint64)
    ni = np.zeros(1, np.uint64)
    csid = np.zeros(n + 1, np.uint64)
    f, nf, inv = (np.zeros(n + 1, np.uint64) for _ in range(3))
    nw = O.lib().orc_warp_layout(O.p64(res), O.p64(ke), n, 64, O.p64(warped), O.p64(wi),
                                 O.p64(ni), O.p64(csid), O.p64(f), O.p64(nf), O.p64(inv))
    n_int = int(ni[0])
    cs = col.searcher(nw, n)
    got = cs.search(warped[:nw], wi[:n_int])
    counts = ci.oracle_counts(res, ke)
    colored = [r for r in range(n) if f[r] > 0]
    assert got.shape == (len(colored), ci.num_colors)
    assert np.array_equal(got, counts[colored])


def test_fused_index_to_color_on_device(ci, gpu_index):
    """reads -> index search with key walk -> color lists, chained on one stream."""
    import torch
    say("fused index->color on device")
    idx, col = gpu_index
    chars, ends = ci.reads(2000, 150, seed=5)
    stream = torch.cuda.Stream()
    s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True, stream=stream.cuda_stream)
    cs = col.searcher(len(chars), len(ends), stream=stream.cuda_stream)
    dc = torch.frombuffer(bytearray(chars), dtype=torch.uint8).cuda()
    de = torch.from_numpy(ends.view(np.int64)).cuda()
    dr = torch.empty(len(chars), dtype=torch.int64, device="cuda")
    d_off = torch.empty(len(ends) + 1, dtype=torch.int64, device="cuda")
    cap = len(ends) * ci.num_colors
    d_ids = torch.empty(cap, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    s.search_reads_device(dc.data_ptr(), len(chars), de.data_ptr(), len(ends), dr.data_ptr())
    nids = cs.pseudoalign_device(dr.data_ptr(), s.kmer_ends_device(), len(ends),
                                 S.color_params(0.7), d_off.data_ptr(), d_ids.data_ptr(), cap)
    stream.synchronize()
    off = d_off.cpu().numpy().view(np.uint64)
    ids = d_ids[:nids].cpu().numpy().view(np.uint32)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    exp = threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke), 0.7)
    got = [ids[int(off[r]):int(off[r + 1])].tolist() for r in range(len(ends))]
    assert got == exp
