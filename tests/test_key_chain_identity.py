"""CPU check of the identity the key pass (k_key_runs) relies on.

The key-k-mer walk from an unmarked node v takes v's first edge (the
smallest letter c with v's bit set) to C[c] + rank_c(v)
(IndexSearcher.cuh:61-69). So when a read's k-mer x at v is followed by the
k-mer y = x[1..k-1] c and c is v's first edge, y's node is that first step
and key(x) = key(y); a marked v is its own key. Both are checked here
against the oracle's independent key walks over every consecutive pair of
found k-mers of sampled reads.
"""
import numpy as np

import oracle_py as O
import workloads as W

K = 31


def _bit(words, i):
    return (int(words[i >> 6]) >> (i & 63)) & 1


def test_key_of_first_edge_successor_is_shared(tmp_path):
    root = W.genome_random(7, 8000)
    gens = [root] + [W.genome_mutate(root, s, 0.02) for s in (8, 9, 10, 11)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, K)
    path = str(tmp_path / "k.sbwt")
    b.write(path)
    km = b.colors(np.arange(len(gens), dtype=np.uint32), d=6).key_marks()
    orc = O.OracleSbwt(path)
    bits = [b.bits(c) for c in range(4)]
    lens = W.uniform_lengths(12, 600, 40, 200)
    genome_ends = np.cumsum([len(g) for g in gens]).astype(np.uint64)
    chars, rends, _ = W.reads(13, np.concatenate(gens), genome_ends, lens, 0.1, 0.01, 0.0)
    nodes = orc.index_reads(chars.tobytes(), rends)
    keys = orc.index_reads(chars.tobytes(), rends, key_marks=km)
    code_of = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3,
               ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}
    same = marked = 0
    slot, start = 0, 0
    for e in rends:
        e = int(e)
        n_k = max(e - start - K + 1, 0)
        for i in range(n_k - 1):
            v, y = int(nodes[slot + i]), int(nodes[slot + i + 1])
            if v >= O.INVALID or y >= O.INVALID:
                continue
            if _bit(km, v):
                assert int(keys[slot + i]) == v
                marked += 1
                continue
            first = next((c for c in range(4) if _bit(bits[c], v)), None)
            c = code_of.get(int(chars[start + i + K]))
            if first is not None and first == c:
                assert int(keys[slot + i]) == int(keys[slot + i + 1]), (v, y)
                same += 1
        slot += n_k
        start = e
    print(f"same-key pairs {same}, marked {marked}")
    assert same > 1000 and marked > 100
