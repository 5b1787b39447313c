"""The load-time suffix-group check (sbwt_suffix_groups_consistent), on the CPU.

In an SBWT only the first node of each suffix group carries the group's edges;
every other node has an empty row. The streaming search steps through group
starts, so the library uses a file's starts only when they pass this check.
Pinned here against the reference's fixture index (every node of it starts a
group), builder-made indexes with multi-node groups, and copies of both with
one bit changed so that the check must fail.
"""
import os

import numpy as np

import sbwt_amd as S
import workloads as W
from conftest import FIXTURES
from sbwt_file import PlainMatrix


def _popcount(words):
    return int(sum(bin(int(x)).count("1") for x in words))


def _check(pm):
    return S.suffix_groups_consistent([pm.vector(c) for c in range(4)], pm.vector(4), pm.num_bits)


def test_reference_fixture_is_consistent_and_all_starts():
    pm = PlainMatrix(os.path.join(FIXTURES, "search_test_index.sbwt"))
    assert pm.num_bits == 236
    assert _check(pm)
    # every node of the fixture starts its own group: it cannot exercise steps
    # from non-start nodes, which the builder-made indexes below do
    assert _popcount(pm.vector(4)) == 236


def test_fixture_with_a_start_cleared_is_rejected(tmp_path):
    pm = PlainMatrix(os.path.join(FIXTURES, "search_test_index.sbwt"))
    edges = pm.vector(0) | pm.vector(1) | pm.vector(2) | pm.vector(3)
    # a node > 0 with an edge: clearing its start leaves an edge on a non-start
    node = next(i for i in range(1, pm.num_bits) if (int(edges[i // 64]) >> (i % 64)) & 1)
    pm.set_bit(4, node, 0)
    assert not _check(pm)
    # the root always starts a group: clearing its bit changes nothing
    pm2 = PlainMatrix(os.path.join(FIXTURES, "search_test_index.sbwt"))
    pm2.set_bit(4, 0, 0)
    assert _check(pm2)


def test_builder_index_has_multi_node_groups_and_passes(tmp_path):
    root = W.genome_random(12, 20_000)
    gens = [root] + [W.genome_mutate(root, s, 0.03) for s in (1, 2, 3)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(tmp_path / "g.sbwt")
    b.write(path)
    pm = PlainMatrix(path)
    assert pm.num_bits == b.n
    assert _check(pm)
    starts = pm.vector(4)
    assert _popcount(starts) < b.n  # some groups hold several nodes
    # an edge bit set on a node that is not a group start breaks the check
    non_start = next(i for i in range(1, b.n) if not (int(starts[i // 64]) >> (i % 64)) & 1)
    for c in range(4):
        assert pm.bit(c, non_start) == 0
    pm.set_bit(2, non_start, 1)
    assert not _check(pm)


def test_bits_past_the_last_node_are_ignored():
    n = 70
    acgt = [np.zeros(2, np.uint64) for _ in range(4)]
    starts = np.zeros(2, np.uint64)
    acgt[1][1] = np.uint64(1 << 20)  # node 84 >= n: outside the SBWT
    assert S.suffix_groups_consistent(acgt, starts, n)
    acgt[1][1] = np.uint64(1 << 3)   # node 67 < n, not a start
    assert not S.suffix_groups_consistent(acgt, starts, n)
    starts[1] = np.uint64(1 << 3)
    assert S.suffix_groups_consistent(acgt, starts, n)
