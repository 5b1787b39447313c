"""A real index of more than 2^32 nodes: the C3 recipe (genomes of 2-5 Mbp from
20 seeded roots, 1-5% substitutions, one color each, d = 20) at 1100 genomes,
4.46 G nodes. It is built on the GPU, written as .sbwt / .tcolors files and
read back through the library's own loaders, so the wide (64-bit node index)
kernels run on an index of that size:
  - both search algorithms (and adaptive work items) against the oracle on
    100,000 reads of U[50, 300] bp (10% random, 1% errors);
  - the key-k-mer move, the color stage (per read and work items) and the fused
    color table (key ranks, 1100 colors) against the oracle on 20,000 reads.
"""
import ctypes as C

import numpy as np
import pytest

import configs
import oracle_py as O
import sbwt_amd as S
import workloads as W
from colorlib import kmer_ends_of, read_stats, threshold_lists

pytestmark = pytest.mark.gpu
GENOMES = 1100


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def wide(tmp_path_factory):
    codes, ends = configs.c3_genomes(n_genomes=GENOMES)
    say(f"{GENOMES} genomes, {len(codes)} bases: GPU build")
    sb, col, st = W.gpu_build(codes, ends, 31, True, np.arange(GENOMES, dtype=np.uint32), 20)
    say(f"built: {st}")
    assert st["nodes"] >= 1 << 32
    d = tmp_path_factory.mktemp("wide")
    paths = (str(d / "wide.sbwt"), str(d / "wide.tcolors"))
    sb.write(paths[0])
    col.write(paths[1])
    key_marks = np.concatenate([col.key_marks(), np.zeros(8, np.uint64)])
    del sb, col
    say("files written")
    return codes, ends, paths, key_marks, st


def _reads(codes, ends, n, seed):
    lens = W.uniform_lengths(seed, n, 50, 300)
    chars, rends, _ = W.reads(seed, codes, ends, lens, random_frac=0.1, err_rate=0.01)
    return chars, rends


def test_wide_index_search_equals_oracle(wide):
    codes, ends, (sbwt_path, _), _, st = wide
    idx = S.Index.load(sbwt_path)
    assert idx.wide and idx.num_nodes == st["nodes"]
    assert idx.has_suffix_groups and not idx.suffix_groups_rejected
    chars, rends = _reads(codes, ends, 100_000, 41)
    say(f"oracle: {len(chars)} chars")
    ref = O.OracleSbwt(sbwt_path).index_reads(chars, rends, threads=16)
    assert np.count_nonzero(ref < np.uint64(idx.num_nodes)) > len(ref) // 2
    assert np.count_nonzero(ref >= np.uint64(1 << 32)) > 0 and \
        np.count_nonzero(ref[ref < np.uint64(idx.num_nodes)] >= np.uint64(1 << 32)) > 0
    s = idx.searcher(len(chars), len(rends))
    for algo in (S.ALGO_STREAMING, S.ALGO_PER_KMER, S.ALGO_STREAMING_ADAPTIVE):
        say(f"wide search algo={algo}")
        s.set_algorithm(algo)
        assert np.array_equal(s.search_reads(chars, rends), ref), algo


def test_wide_key_moves_and_colors_equal_oracle(wide):
    codes, ends, (sbwt_path, colors_path), key_marks, st = wide
    idx = S.Index.load(sbwt_path, colors_path)
    assert idx.wide and idx.has_key_marks
    chars, rends = _reads(codes, ends, 20_000, 43)
    say("oracle key moves")
    orc = O.OracleSbwt(sbwt_path)
    keys = orc.index_reads(chars, rends, key_marks=key_marks, threads=16)
    s = idx.searcher(len(chars), len(rends), move_to_key_kmer=True)
    for algo in (S.ALGO_STREAMING, S.ALGO_PER_KMER):
        s.set_algorithm(algo)
        assert np.array_equal(s.search_reads(chars, rends), keys), algo
    say("oracle colors")
    oc = O.Colors()
    assert O.lib().orc_colors_load(colors_path.encode(), C.byref(oc)) == 0
    nc = int(oc.num_colors)
    assert nc == GENOMES
    ke = kmer_ends_of(rends)
    n = len(ke)
    counts = np.zeros(n * nc, np.uint64)
    O.lib().orc_color_counts(C.byref(oc), O.p64(keys), O.p64(ke), n, O.p64(counts), 16)
    counts = counts.reshape(n, nc)
    stats = read_stats(keys, ke)
    exp = threshold_lists(counts, stats, 0.7, first_read_quirk=True)
    col = S.ColorIndex(colors_path)
    params = S.color_params(0.7, False, False, True)
    for items in (0, 256):
        say(f"wide colors items={items}")
        cs = col.searcher(len(keys) + 64, n)
        cs.set_work_items(items)
        lists, _, _, _ = cs.pseudoalign(keys, ke, params)
        assert lists == exp, items
    assert sum(len(x) for x in exp) > 0
    say("fused color table (key ranks)")
    assert idx.build_color_table(col) == 2
    s.set_color_output(True)
    entries = s.search_reads(chars, rends)
    cs = col.searcher(len(entries) + 64, n)
    cs.set_input(idx)
    lists, _, _, _ = cs.pseudoalign(entries, ke, params)
    assert lists == exp
