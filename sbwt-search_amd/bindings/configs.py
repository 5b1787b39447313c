"""The seeded synthetic workloads the search path is measured and tested on.

Each builder returns plain numpy data; nothing here touches the GPU.

  C1  1 Mbp uniform genome (seed 0), SBWT k=31; 10,000 reads x 150 bp, even reads
      sampled (uniform position and strand, 1% substitutions), odd reads uniform
      random, then 0.1% of bases -> 'N' (reads seed 1).
  C2  root 2 Mbp (seed 0); 50 genomes = root with 2% substitutions and 0.1%
      indels of length 1-10 (seeds 1..50); union SBWT k=31 (~95 M nodes);
      10,000,000 reads x 150 bp: 95% from a random genome/strand with 1%
      substitutions, 5% uniform random (reads seed 1).
"""
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
