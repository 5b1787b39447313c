"""The GPU builder (bld_gpu_build) against the host builder: the .sbwt and
.tcolors files they write are byte-identical, over inputs that reach every
branch of the construction: reverse complements on and off, k from 2 to 31,
N codes, sequences shorter than k, empty sequences, colors given out of order,
short reads (many dummy nodes) and cyclic sequences (chains of the successor
graph with no first node, whose key marks the host sets in node order).

The host builder's files are pinned by the reference's own index fixtures
through the search tests; equality carries that to the GPU builder.
"""
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def _genomes(seed, n, length, sub=0.02, with_n=True):
    root = W.genome_random(seed, length)
    gens = [W.genome_mutate(root, seed * 100 + i, sub, 0.001, 5) for i in range(n)]
    if with_n:
        rng = np.random.default_rng(seed)
        for g in gens[::2]:
            g[rng.integers(0, len(g), size=max(1, len(g) // 2000))] = 4
    return gens


def _cyclic(seed, length, k):
    x = W.genome_random(seed, length)
    return np.concatenate([x, x[:k - 1]])  # its k-mers form one cycle


CASES = {
    "k31-rc-colors": dict(seqs=lambda: _genomes(1, 6, 20_000) + [np.zeros(0, np.uint8),
                                                                W.genome_random(5, 20)],
                          k=31, rc=True, colors=[3, 1, 4, 1, 0, 2, 2, 5], d=5),
    "k15-fw-only": dict(seqs=lambda: _genomes(2, 4, 15_000), k=15, rc=False,
                        colors=[0, 1, 2, 3], d=7),
    "k5-dense": dict(seqs=lambda: _genomes(3, 5, 4000), k=5, rc=True, colors=[0, 1, 2, 3, 4],
                     d=3),
    "k2": dict(seqs=lambda: _genomes(4, 3, 500), k=2, rc=True, colors=[0, 0, 1], d=2),
    "reads": dict(seqs=lambda: [W.genome_mutate(W.genome_random(6, 50_000)[s:s + 60], 700 + i,
                                                0.01)
                                for i, s in enumerate(range(0, 49_000, 25))],
                  k=31, rc=True, colors=None, d=20),
    "cycles": dict(seqs=lambda: [_cyclic(7, 300, 31), _cyclic(8, 1000, 31), W.genome_random(9, 5000)],
                   k=31, rc=True, colors=[0, 1, 1], d=4),
    "cycles-d1": dict(seqs=lambda: [_cyclic(10, 200, 21)], k=21, rc=False, colors=[0], d=1),
    "short-only": dict(seqs=lambda: [W.genome_random(11, 10), np.zeros(0, np.uint8)], k=31,
                       rc=True, colors=[0, 1], d=20),
}


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_build_files_equal_host_build(tmp_path, name):
    c = CASES[name]
    seqs = c["seqs"]()
    codes, ends = W.concat(seqs) if seqs else (np.zeros(0, np.uint8), np.zeros(0, np.uint64))
    colors = None if c["colors"] is None else np.array(c["colors"][:len(seqs)], np.uint32)
    say(f"{name}: {len(seqs)} sequences, {len(codes)} bases, k={c['k']} rc={c['rc']}")
    host = W.BuiltSbwt(codes, ends, c["k"], c["rc"])
    gpu, gcol, st = W.gpu_build(codes, ends, c["k"], c["rc"], colors, c["d"], keep_kmers=True)
    assert (gpu.n, gpu.n_kmers) == (host.n, host.n_kmers)
    assert st["nodes"] == host.n and st["kmers"] == host.n_kmers
    for ch in range(5):
        assert np.array_equal(gpu.bits(ch), host.bits(ch)), ch
    host.write(str(tmp_path / "h.sbwt"))
    gpu.write(str(tmp_path / "g.sbwt"))
    assert (tmp_path / "h.sbwt").read_bytes() == (tmp_path / "g.sbwt").read_bytes()
    if host.n_kmers:
        keys = np.array([W.kmer_key(codes[e - c["k"]:e]) for e in ends[:3] if e >= c["k"]
                         and codes[e - c["k"]:e].max() < 4], np.uint64)
        assert np.array_equal(gpu.lookup(keys), host.lookup(keys))
    if colors is None:
        assert gcol is None
        return
    hcol = host.colors(colors, c["d"])
    assert (gcol.num_colors, gcol.num_sets, gcol.num_dense) == \
        (hcol.num_colors, hcol.num_sets, hcol.num_dense)
    assert np.array_equal(gcol.key_marks(), hcol.key_marks())
    hcol.write(str(tmp_path / "h.tcolors"))
    gcol.write(str(tmp_path / "g.tcolors"))
    assert (tmp_path / "h.tcolors").read_bytes() == (tmp_path / "g.tcolors").read_bytes()
    if name.startswith("cycles"):
        assert st["cycle_nodes"] > 0


def test_gpu_build_without_kmer_list_refuses_lookup():
    codes, ends = W.concat(_genomes(12, 2, 3000))
    gpu, _, _ = W.gpu_build(codes, ends, 31)
    with pytest.raises(RuntimeError):
        gpu.lookup(np.zeros(1, np.uint64))


def test_gpu_build_rejects_k32():
    codes, ends = W.concat([W.genome_random(13, 100)])
    with pytest.raises(RuntimeError, match="k in"):
        W.gpu_build(codes, ends, 32)


def test_gpu_build_c2_scale_equals_host(tmp_path):
    """50 genomes x 2 Mbp with indels (the plain search index, 100 Mbp): the
    GPU build equals the host build, and its stage times are reported."""
    import configs
    codes, ends = configs.c2_genomes()
    say(f"C2 genomes: {len(codes)} bases")
    gpu, gcol, st = W.gpu_build(codes, ends, 31, True, np.arange(len(ends), dtype=np.uint32), 20)
    say(f"GPU build: {st}")
    host = W.BuiltSbwt(codes, ends, 31)
    assert gpu.n == host.n
    for ch in range(5):
        assert np.array_equal(gpu.bits(ch), host.bits(ch)), ch
    hcol = host.colors(np.arange(len(ends), dtype=np.uint32), 20)
    assert np.array_equal(gcol.key_marks(), hcol.key_marks())
    gcol.write(str(tmp_path / "g.tcolors"))
    hcol.write(str(tmp_path / "h.tcolors"))
    assert (tmp_path / "h.tcolors").read_bytes() == (tmp_path / "g.tcolors").read_bytes()
