"""The color stage's census (sbwt_color_searcher_set_census): the exact count of
key-set entries, key-mark entries and payload lines that one pseudoalign call
reads, against the same quantities worked out on the host from the oracle's
view of the color index (.tcolors loaded by the oracle's reader).

The color kernel reads one key set per run of consecutive equal found results
(consecutive k-mers that walked to the same key k-mer), the runs being cut at the
read's (or work item's) groups of 256 k-mers; key-node input also reads one key-
mark entry per run; each set's payload spans the 128-byte lines of its bits
(dense: one bit per color; sparse: one int_vector element per id).
"""

import numpy as np
import pytest

import sbwt_amd as S
import workloads as W
from colorlib import ColoredIndex, kmer_ends_of

pytestmark = pytest.mark.gpu
GROUP = 256  # k-mers decoded together by one wave


def say(msg):
    print(f"[case] {msg}", flush=True)


def _intv(v, i):
    """Element i of an oracle int_vector (LSB-first packing)."""
    words = np.ctypeslib.as_array(v.data, shape=(int(v.size) * int(v.width) // 64 + 2,))
    bit = int(i) * int(v.width)
    w, off = bit >> 6, bit & 63
    x = int(words[w]) >> off
    if off + int(v.width) > 64:
        x |= int(words[w + 1]) << (64 - off)
    return x & ((1 << int(v.width)) - 1)


def _bit(b, i):
    words = np.ctypeslib.as_array(b.data, shape=((int(b.bits) + 63) // 64,))
    return (int(words[i >> 6]) >> (i & 63)) & 1


def _lines(b0, b1):
    return ((b1 - 1) >> 10) - (b0 >> 10) + 1 if b1 > b0 else 0


def expected_census(ci, keys, ke, item_kmers=0):
    """{set_lines, mark_lines, payload_lines, kmers} of one call on key-node input."""
    oc = ci.oc
    km = ci.key_marks
    bits = np.unpackbits(km.view(np.uint8), bitorder="little").astype(np.int64)
    ranks = np.concatenate([[0], np.cumsum(bits)])
    n_nodes = int(ci.sbwt.n)
    n_sets = int(oc.num_color_sets)
    dense_rank = np.concatenate([[0], np.cumsum(
        [_bit(oc.is_dense_marks, c) for c in range(n_sets)])]).astype(np.int64)
    swidth = int(oc.sparse_arrays.width)
    set_lines = payload = 0
    start = 0
    for e in ke.astype(np.int64):
        segs = [(start, e)] if not item_kmers else \
            [(a, min(e, a + item_kmers)) for a in range(start, max(e, start + 1), item_kmers)]
        for a, b in segs:
            for g0 in range(a, b, GROUP):
                v = keys[g0:min(b, g0 + GROUP)]
                for i, x in enumerate(v):
                    x = int(x)
                    if x >= n_nodes or (i and int(v[i - 1]) == x):
                        continue
                    set_lines += 1
                    cs = _intv(oc.color_set_idxs, ranks[x])
                    if cs >= n_sets:
                        continue
                    if _bit(oc.is_dense_marks, cs):
                        d = int(dense_rank[cs])
                        s0 = _intv(oc.dense_arrays_intervals, d)
                        s1 = _intv(oc.dense_arrays_intervals, d + 1)
                        payload += _lines(s0, s1)
                    else:
                        q = cs - int(dense_rank[cs])
                        s0 = _intv(oc.sparse_arrays_intervals, q)
                        s1 = _intv(oc.sparse_arrays_intervals, q + 1)
                        payload += _lines(s0 * swidth, s1 * swidth)
        start = int(e)
    return {"set_lines": set_lines, "mark_lines": set_lines, "payload_lines": payload,
            "kmers": int(ke[-1]) if len(ke) else 0}


@pytest.fixture(scope="module")
def colored(tmp_path_factory):
    ci = ColoredIndex(tmp_path_factory.mktemp("census"), n_genomes=70, length=8_000, d=7,
                      sub=0.03, seed=9)
    lens = np.concatenate([W.uniform_lengths(71, 500, 31, 320), np.full(40, 1400, np.uint32),
                           np.array([0, 5, 31, 32, 287, 288], np.uint32)])
    lens = lens[W.permutation(72, len(lens)).astype(np.int64)]
    chars, ends = ci.reads(len(lens), lens, seed=73, random_frac=0.1, err=0.01)
    keys = ci.index_results(chars, ends)
    return ci, chars, ends, keys, kmer_ends_of(ends)


@pytest.mark.parametrize("items", [0, 64, 1024])
def test_color_census_equals_host_count(colored, items):
    ci, _, _, keys, ke = colored
    say(f"color census, key-node input, work items {items}")
    col = S.ColorIndex(ci.colors_path)
    cs = col.searcher(len(keys) + 64, len(ke))
    cs.set_work_items(items)
    cs.set_census(True)
    cs.pseudoalign(keys, ke, S.color_params(0.7))
    got = cs.last_census()
    exp = expected_census(ci, keys, ke, items)
    assert exp["set_lines"] > 500 and exp["payload_lines"] > 0
    assert got == exp
    cs.set_census(False)
    lists_off, _, _, _ = cs.pseudoalign(keys, ke, S.color_params(0.7))
    cs.set_census(True)
    lists_on, _, _, _ = cs.pseudoalign(keys, ke, S.color_params(0.7))
    assert lists_on == lists_off  # the census leaves results unchanged


def test_color_census_rank_input(colored):
    """The fused form's key ranks read the same sets and payload lines, no marks."""
    ci, chars, ends, keys, ke = colored
    say("color census, key-rank input")
    idx = S.Index.load(ci.sbwt_path, ci.colors_path)
    col = S.ColorIndex(ci.colors_path)
    assert idx.build_color_table(col) == 2  # 70 colors: key ranks
    s = idx.searcher(len(chars), len(ends), move_to_key_kmer=True)
    s.set_color_output(True)
    ranks = s.search_reads(chars, ends)
    cs = col.searcher(len(keys) + 64, len(ke))
    cs.set_input(idx)
    cs.set_census(True)
    cs.pseudoalign(ranks, ke, S.color_params(0.7))
    got = cs.last_census()
    exp = expected_census(ci, keys, ke)
    exp["mark_lines"] = 0
    assert got == exp
