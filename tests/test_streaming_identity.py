"""CPU check of the identity the streaming search kernel relies on.

For a found k-mer x at node v and a letter c, the k-mer y = x[1..k-1] c is
  C[c] + rank_c(v)   if v carries edge c,
  absent             if v carries another edge,
  else decided by the suffix-group start gs <= v the same way (absent if gs = v),
which must equal an independent search of y (IndexSearcher.cuh:22-72).
"""
import numpy as np

import oracle_py as O
import workloads as W

K = 31


def _bit(words, i):
    return (int(words[i >> 6]) >> (i & 63)) & 1


def _rank(words, i):
    pc = np.unpackbits(words.view(np.uint8), bitorder="little")
    return int(pc[:i].sum())


def test_stream_step_equals_fresh_search(tmp_path):
    root = W.genome_random(3, 6000)
    gens = [root] + [W.genome_mutate(root, s, 0.03) for s in (4, 5, 6)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, K)
    path = str(tmp_path / "s.sbwt")
    b.write(path)
    orc = O.OracleSbwt(path)
    bits = [b.bits(c) for c in range(4)]
    sgs = b.bits(4)
    cm = orc.c_map()
    prefix = [np.concatenate([[0], np.cumsum(np.unpackbits(x.view(np.uint8), bitorder="little"))])
              for x in bits]

    def rank(c, i):
        return int(prefix[c][i])

    def group_start(v):
        while not _bit(sgs, v):
            v -= 1
        return v

    def stream(v, c):
        if _bit(bits[c], v):
            return cm[c] + rank(c, v)
        if any(_bit(bits[cc], v) for cc in range(4) if cc != c):
            return int(O.NOT_FOUND)
        gs = group_start(v)
        if gs == v or not _bit(bits[c], gs):
            return int(O.NOT_FOUND)
        return cm[c] + rank(c, gs)

    rng = np.random.default_rng(1)
    checked = non_leader = 0
    for g in gens:
        for st in rng.integers(0, len(g) - K + 1, 150):
            x = g[st:st + K]
            v = orc.search_codes(x)
            assert v != O.NOT_FOUND
            non_leader += not _bit(sgs, v)
            for c in range(4):
                y = np.concatenate([x[1:], [c]]).astype(np.uint8)
                assert stream(v, c) == orc.search_codes(y)
                checked += 1
    assert checked == 4 * 150 * len(gens)
    assert non_leader > 0  # the group-start path is exercised
