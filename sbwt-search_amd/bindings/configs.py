"""The seeded synthetic workloads the search path is measured and tested on.

Each builder returns plain numpy data; nothing here touches the GPU.

  C1  1 Mbp uniform genome (seed 0), SBWT k=31; 10,000 reads x 150 bp, even reads
      sampled (uniform position and strand, 1% substitutions), odd reads uniform
      random, then 0.1% of bases -> 'N' (reads seed 1).
  C2  root 2 Mbp (seed 0); 50 genomes = root with 2% substitutions and 0.1%
      indels of length 1-10 (seeds 1..50); union SBWT k=31 (~95 M nodes);
      10,000,000 reads x 150 bp: 95% from a random genome/strand with 1%
      substitutions, 5% uniform random (reads seed 1).
  C3  color search: genomes derived from seeded uniform roots with 1-5%
      substitutions (rate uniform per genome), one color per genome, key k-mers
      every d = 20 on non-branching paths; reads of length U[50, 300] drawn from
      random genomes with 1% error (reads seed 3). Full scale is 1,000 genomes of
      2-5 Mbp from 20 roots; smaller genome counts/lengths keep the same recipe.
  C4  batched form: the C2 genomes and their color index (50 colors, d = 20);
      1,000,000 reads with log-normal lengths (median 2,000 bp, sigma 1.2,
      clipped to [31, 100,000]) and 4,000,000 reads x 150 bp, interleaved in a
      seeded random order (reads seed 4); sampled as in C2.
  C5  serving index: 200 genomes of 5 Mbp, genome g = root g mod 10 with 3%
      substitutions, one color per genome; requests of N log-uniform in [1, 10^5]
      reads of length U[75, 250].
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import workloads as W

K = 31


class Workload:
    def __init__(self, name, description, genomes, genome_ends, chars, read_ends):
        self.name = name
        self.description = description
        self.genomes = genomes
        self.genome_ends = genome_ends
        self.chars = chars
        self.read_ends = read_ends

    @property
    def n_reads(self):
        return len(self.read_ends)

    @property
    def n_chars(self):
        return len(self.chars)

    def n_kmers(self, k=K):
        lens = np.diff(np.concatenate([[0], self.read_ends.astype(np.int64)]))
        return int(np.maximum(lens - k + 1, 0).sum())


def c1(n_reads=10_000):
    g = W.genome_random(0, 1_000_000)
    codes, ends = W.concat([g])
    chars, rends, _ = W.reads(1, codes, ends, np.full(n_reads, 150, np.uint32),
                              err_rate=0.01, n_rate=0.001, alternate=True)
    return Workload("C1", "plain index search: 1 Mbp uniform genome (seed 0), SBWT k=31; "
                    f"{n_reads} reads x 150 bp, half sampled with 1% substitutions, half "
                    "random, 0.1% N", codes, ends, chars, rends)


def c2_genomes(n_genomes=50, length=2_000_000):
    root = W.genome_random(0, length)
    gens = [W.genome_mutate(root, s, 0.02, 0.001, 10) for s in range(1, n_genomes + 1)]
    return W.concat(gens)


def c2(n_reads=10_000_000, n_genomes=50, length=2_000_000):
    codes, ends = c2_genomes(n_genomes, length)
    chars, rends, _ = W.reads(1, codes, ends, np.full(n_reads, 150, np.uint32),
                              random_frac=0.05, err_rate=0.01)
    return Workload("C2", f"plain index search at scale: {n_genomes} genomes x {length} bp from "
                    "one seeded root (2% substitutions, 0.1% indels of 1-10 bp), union SBWT "
                    f"k=31; {n_reads} reads x 150 bp (95% sampled with 1% substitutions, 5% "
                    "random)", codes, ends, chars, rends)


def c3_genomes(n_genomes=1000, n_roots=20, min_len=2_000_000, max_len=5_000_000):
    roots = [W.genome_random(100 + i, max_len) for i in range(n_roots)]

    def genome(g):  # seeded per genome, so the threads' order does not matter
        length = int(min_len + W.lib().bld_draw(3, 11, g) % (max_len - min_len + 1))
        rate = 0.01 + 0.04 * (W.lib().bld_draw(3, 12, g) >> 11) * 2.0 ** -53
        return W.genome_mutate(roots[g % n_roots][:length], 1000 + g, rate)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        gens = list(ex.map(genome, range(n_genomes)))
    return W.concat(gens)


def c3(n_reads=5_000_000, n_genomes=1000, n_roots=20, min_len=2_000_000, max_len=5_000_000):
    codes, ends = c3_genomes(n_genomes, n_roots, min_len, max_len)
    lens = W.uniform_lengths(3, n_reads, 50, 300)
    chars, rends, _ = W.reads(3, codes, ends, lens, random_frac=0.0, err_rate=0.01)
    return Workload("C3", f"color search: {n_genomes} genomes of {min_len}-{max_len} bp from "
                    f"{n_roots} seeded roots (1-5% substitutions), one color each, d=20; "
                    f"{n_reads} reads of U[50,300] bp with 1% error", codes, ends, chars, rends)


def c5_genomes(n_genomes=200, n_roots=10, length=5_000_000, rate=0.03):
    """Serving index: genome g is root g mod n_roots with `rate` substitutions."""
    roots = [W.genome_random(200 + i, length) for i in range(n_roots)]
    gens = [W.genome_mutate(roots[g % n_roots], 2000 + g, rate) for g in range(n_genomes)]
    return W.concat(gens)


def c5_index(n_reads=10_000_000, n_genomes=200, n_roots=10, length=5_000_000):
    """The serving index (C5 recipe) searched with a batch of reads of U[75, 250] bp
    (1% error): an index whose node lines are far larger than the Infinity Cache."""
    codes, ends = c5_genomes(n_genomes, n_roots, length)
    lens = W.uniform_lengths(5, n_reads, 75, 250)
    chars, rends, _ = W.reads(5, codes, ends, lens, random_frac=0.0, err_rate=0.01)
    return Workload("C5", f"index search on the serving index: {n_genomes} genomes x {length} bp "
                    f"from {n_roots} seeded roots (3% substitutions), SBWT k=31; {n_reads} reads "
                    "of U[75,250] bp with 1% error", codes, ends, chars, rends)


def c4(n_long=1_000_000, n_short=4_000_000, n_genomes=50, length=2_000_000):
    codes, ends = c2_genomes(n_genomes, length)
    long_lens = W.lognormal_lengths(4, n_long, 2000.0, 1.2, 31, 100_000)
    lens = np.concatenate([long_lens, np.full(n_short, 150, np.uint32)])
    lens = np.ascontiguousarray(lens[W.permutation(4, len(lens)).astype(np.int64)])
    chars, rends, _ = W.reads(4, codes, ends, lens, random_frac=0.05, err_rate=0.01)
    return Workload("C4", f"batched form: C2 genomes + color index (50 colors, d=20); "
                    f"{n_long} log-normal reads (median 2 kbp, sigma 1.2, [31, 100 kbp]) "
                    f"interleaved with {n_short} reads x 150 bp", codes, ends, chars, rends)

