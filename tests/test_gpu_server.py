"""Serving loop on the MI355X: every request's result equals its result alone."""
import numpy as np
import pytest

import sbwt_amd as S
from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def served(tmp_path_factory):
    ci = ColoredIndex(tmp_path_factory.mktemp("srv"), n_genomes=6, length=20_000, d=8)
    return ci, S.Index.load(ci.sbwt_path, ci.colors_path), S.ColorIndex(ci.colors_path)


def _requests(ci, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        n_reads = int(np.exp(rng.uniform(0, np.log(3000))))
        lens = rng.integers(75, 251, n_reads)
        out.append(ci.reads(n_reads, lens, seed=seed * 1000 + i))
    return out


def test_served_results_equal_alone(served):
    ci, idx, col = served
    say("serving: mix vs alone")
    reqs = _requests(ci, 24, seed=1)
    params = S.color_params(0.7)
    srv = S.Server(idx, col, slots=4, max_chars_per_slot=200_000, max_reads_per_slot=2000,
                   move_to_key_kmer=True, params=params)
    assert srv.slots >= 2
    ids = [srv.submit(c, e) for c, e in reqs]
    got = [srv.wait(i) for i in ids]
    assert srv.peak_in_flight() >= 2  # requests overlapped on the device
    alone = idx.searcher(200_000 * 4, 50_000, move_to_key_kmer=True)
    cs = col.searcher(200_000 * 4, 50_000)
    for (chars, ends), g in zip(reqs, got):
        res = alone.search_reads(chars, ends)
        assert np.array_equal(g["kmers"], res)
        assert np.array_equal(res, ci.index_results(chars, ends))
        ke = kmer_ends_of(ends)
        lists, _, _, _ = cs.pseudoalign(res, ke, params)
        assert g["color_lists"] == lists
        assert lists == threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke), 0.7)
        assert g["submit_s"] <= g["start_s"] <= g["end_s"]


@pytest.mark.parametrize("kmer_results", [True, False], ids=["kmer-results", "lists-only"])
def test_reads_longer_than_a_slot_are_searched_in_pieces(served, kmer_results):
    """Slots of 1000 chars: reads of up to 12 kbp are searched in pieces that
    overlap by k - 1 chars, their color counts summed over the pieces; k-mer
    results and color lists equal the oracle's for the whole reads."""
    ci, idx, col = served
    say(f"serving: reads longer than a slot, kmer_results={kmer_results}")
    rng = np.random.default_rng(21)
    reqs = []
    for i in range(6):
        n = int(rng.integers(3, 40))
        lens = rng.integers(20, 300, n)
        lens[rng.integers(0, n, 3)] = rng.integers(1001, 12_000, 3)
        reqs.append(ci.reads(n, lens, seed=500 + i))
    params = S.color_params(0.7)
    srv = S.Server(idx, col, slots=2, max_chars_per_slot=1000, max_reads_per_slot=64,
                   move_to_key_kmer=True, params=params, kmer_results=kmer_results)
    ids = [srv.submit(c, e) for c, e in reqs]
    n_long = 0
    for (chars, ends), rid in zip(reqs, ids):
        g = srv.wait(rid)
        res = ci.index_results(chars, ends)
        if kmer_results:
            assert np.array_equal(g["kmers"], res)
        ke = kmer_ends_of(ends)
        assert g["color_lists"] == threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke),
                                                   0.7)
        n_long += int(np.count_nonzero(np.diff(np.concatenate([[0], ends.astype(np.int64)]))
                                       > 1000))
    assert n_long >= 6


def test_malformed_request_does_not_disturb_others(served):
    ci, idx, col = served
    say("serving: malformed request in the mix")
    reqs = _requests(ci, 6, seed=2)
    srv = S.Server(idx, col, slots=3, max_chars_per_slot=200_000, max_reads_per_slot=2000,
                   move_to_key_kmer=True, params=S.color_params(0.7))
    ids = [srv.submit(c, e) for c, e in reqs[:3]]
    bad_chars, bad_ends = reqs[3]
    bad = srv.submit(bad_chars, bad_ends[::-1].copy())  # decreasing read ends
    ids += [srv.submit(c, e) for c, e in reqs[4:]]
    with pytest.raises(S.SbwtError):
        srv.wait(bad)
    alone = idx.searcher(200_000 * 4, 50_000, move_to_key_kmer=True)
    for rid, (chars, ends) in zip(ids, reqs[:3] + reqs[4:]):
        assert np.array_equal(srv.wait(rid)["kmers"], alone.search_reads(chars, ends))


def test_graph_replays_equal_eager(served):
    """Small batches replay a HIP graph captured once per slot and size bucket.
    The same requests served several times (so every graph is replayed, with
    direct searches on the device in between) equal a server without graphs."""
    ci, idx, col = served
    say("serving: graph replays vs eager")
    rng = np.random.default_rng(9)
    reqs = []
    for i in range(12):
        n_reads = int(rng.integers(1, 400))
        reqs.append(ci.reads(n_reads, rng.integers(20, 251, n_reads), seed=9000 + i))
    params = S.color_params(0.7)
    kw = dict(slots=2, max_chars_per_slot=1 << 20, max_reads_per_slot=1 << 14,
              move_to_key_kmer=True, params=params)
    srv_g = S.Server(idx, col, graphs=True, **kw)
    srv_e = S.Server(idx, col, graphs=False, **kw)
    other = idx.searcher(1 << 20, 1 << 14, move_to_key_kmer=True)
    for rep in range(3):
        for chars, ends in reqs:
            g = srv_g.wait(srv_g.submit(chars, ends))
            e = srv_e.wait(srv_e.submit(chars, ends))
            assert np.array_equal(g["kmers"], e["kmers"]), rep
            assert g["color_lists"] == e["color_lists"], rep
            assert np.array_equal(g["kmers"], ci.index_results(chars, ends))
            other.search_reads(chars, ends)  # eager device work between replays


def test_replay_load_generator(served):
    """sbwt_server_replay: requests submitted on a schedule from C++; outputs of the
    served form equal the one-slot plain form's request by request (digests), and
    latency runs from the scheduled arrival."""
    import modes
    ci, idx, col = served
    say("serving: replay load generator, served vs plain")
    import workloads as W

    def reads(seed, n):
        lens = W.uniform_lengths(seed, n, 75, 250)
        c, e, _ = W.reads(seed, ci.codes, ci.ends, lens, random_frac=0.05, err_rate=0.01)
        return c, e

    out = modes.run_serve(S, idx, col, reads, 400.0, 40, 3000, 3, S.color_params(0.7),
                          slots=4, log=say)
    assert out["outputs_equal"]
    for form in ("served", "served_no_graphs", "plain"):
        r = out[form]
        assert 0 <= r["p50_ms"] <= r["p99_ms"] <= r["max_ms"]
        assert r["submit_lag_max_ms"] >= 0 and r["reads_per_s"] > 0
    assert out["plain"]["slots"] == 1 and out["served"]["slots"] >= 2
    # color lists only: the forms still agree request by request
    out2 = modes.run_serve(S, idx, col, reads, 400.0, 20, 2000, 4, S.color_params(0.7),
                           slots=4, log=say, kmer_results=False)
    assert out2["outputs_equal"] and out2["returns"] == "color lists"
    # one request's digest matches its own search alone
    srv = S.Server(idx, col, slots=2, max_chars_per_slot=1 << 20, max_reads_per_slot=3000,
                   move_to_key_kmer=True, params=S.color_params(0.7))
    c, e = reads(77, 500)
    r1 = srv.replay(c, e, np.array([0, len(e)], np.uint64), np.zeros(1))
    r2 = srv.replay(c, e, np.array([0, len(e)], np.uint64), np.zeros(1))
    assert r1["digest"][0] == r2["digest"][0] != 0
    rid = srv.submit(c.tobytes(), e)
    g = srv.wait(rid)
    assert np.array_equal(g["kmers"], ci.index_results(c.tobytes(), e))
    srv.close()
