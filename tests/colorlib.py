"""Helpers of the color tests: small colored indexes with brute-force color sets."""
import ctypes as C

import numpy as np

import oracle_py as O
import workloads as W

K = 31


def rc_codes(x):
    return (3 - x)[::-1]


class ColoredIndex:
    """Genomes (one color each unless `colors` says otherwise), SBWT + .tcolors on disk,
    the oracle's view of both, and brute-force k-mer -> color set."""

    def __init__(self, tmp, n_genomes=8, length=30_000, sub=0.02, d=5, seed=1, colors=None):
        root = W.genome_random(seed, length)
        self.genomes = [root] + [W.genome_mutate(root, seed * 100 + s, sub)
                                 for s in range(1, n_genomes)]
        self.codes, self.ends = W.concat(self.genomes)
        self.color_of = np.asarray(colors if colors is not None else np.arange(n_genomes),
                                   np.uint32)
        self.sbwt = W.BuiltSbwt(self.codes, self.ends, K)
        self.colors = self.sbwt.colors(self.color_of, d=d)
        self.sbwt_path = str(tmp / "c.sbwt")
        self.colors_path = str(tmp / "c.tcolors")
        self.sbwt.write(self.sbwt_path)
        self.colors.write(self.colors_path)
        self.orc = O.OracleSbwt(self.sbwt_path)
        self.key_marks = np.concatenate([self.colors.key_marks(), np.zeros(8, np.uint64)])
        self.oc = O.Colors()
        if O.lib().orc_colors_load(self.colors_path.encode(), C.byref(self.oc)):
            raise RuntimeError("oracle could not load the color index")
        self.num_colors = int(self.oc.num_colors)
        self._kmer_colors = None

    def kmer_colors(self):
        """colex key -> frozenset of colors of genomes holding the k-mer or its revcomp."""
        if self._kmer_colors is None:
            m = {}
            for g, seq in enumerate(self.genomes):
                col = int(self.color_of[g])
                keys = set()
                for i in range(len(seq) - K + 1):
                    x = seq[i:i + K]
                    keys.add(W.kmer_key(x))
                    keys.add(W.kmer_key(rc_codes(x)))
                for k in keys:
                    m.setdefault(k, set()).add(col)
            self._kmer_colors = {k: frozenset(v) for k, v in m.items()}
        return self._kmer_colors

    def reads(self, n, length, seed, random_frac=0.2, err=0.01, n_rate=0.002):
        lens = np.full(n, length, np.uint32) if np.isscalar(length) else \
            np.asarray(length, np.uint32)
        chars, ends, _ = W.reads(seed, self.codes, self.ends, lens, random_frac, err, n_rate)
        return chars.tobytes(), ends

    def index_results(self, chars, ends):
        """Oracle index results with the key-k-mer move (the colors stage input)."""
        return self.orc.index_reads(chars, ends, key_marks=self.key_marks)

    def color_set(self, node):
        ids = np.empty(self.num_colors + 1, np.uint32)
        n = O.lib().orc_color_set_of(C.byref(self.oc), int(node), O.p32(ids))
        return [int(x) for x in ids[:n]]

    def oracle_counts(self, results, kmer_ends):
        n = len(kmer_ends)
        out = np.zeros(max(n, 1) * self.num_colors, np.uint64)
        O.lib().orc_color_counts(C.byref(self.oc), O.p64(np.ascontiguousarray(results, np.uint64)),
                                 O.p64(np.ascontiguousarray(kmer_ends, np.uint64)), n, O.p64(out),
                                 4)
        return out[:n * self.num_colors].reshape(n, self.num_colors)


def kmer_ends_of(read_ends, k=K):
    lens = np.diff(np.concatenate([[0], np.asarray(read_ends, np.int64)]))
    return np.cumsum(np.maximum(lens - k + 1, 0)).astype(np.uint64)


def read_stats(results, kmer_ends):
    out = []
    st = 0
    for e in kmer_ends:
        v = results[st:int(e)]
        inv = int(np.count_nonzero(v == O.INVALID))
        nf = int(np.count_nonzero(v == O.NOT_FOUND))
        out.append((len(v) - inv - nf, nf, inv))
        st = int(e)
    return out


def colored_seq_ids(stats):
    """colored_seq_id of every read (IndexFileParser.cpp:65-74 through the oracle's
    orc_warp_layout): a read's colored slot, or for a read with no found index the
    slot of the next read that has one (= the number of colored reads before it)."""
    n = len(stats)
    res, ends = [], []
    for f, nf, inv in stats:
        res += [0] * f + [int(O.NOT_FOUND)] * nf + [int(O.INVALID)] * inv
        ends.append(len(res))
    res = np.array(res + [0], np.uint64)
    ends = np.array(ends + [0], np.uint64)
    warped = np.zeros(len(res) + 64 * (n + 1), np.uint64)
    wi = np.zeros(n + 2, np.uint64)
    ni = np.zeros(1, np.uint64)
    cs = np.zeros(n + 1, np.uint64)
    f_, nf_, inv_ = (np.zeros(n + 1, np.uint64) for _ in range(3))
    O.lib().orc_warp_layout(O.p64(res), O.p64(ends), n, 64, O.p64(warped), O.p64(wi), O.p64(ni),
                            O.p64(cs), O.p64(f_), O.p64(nf_), O.p64(inv_))
    return [int(x) for x in cs[:n]]


def threshold_lists(counts, stats, tau, inc_nf=False, inc_inv=False, first_read_quirk=False,
                    quirk2=True):
    """Sorted color ids per read: count >= ceil(seq_size * tau), u64 wrap for the quirk.
    Quirk 2 (ContinuousColorResultsPrinter.hpp:201-209): each read is printed with the
    counts of its colored_seq_id slot; a read with no found index therefore takes the
    counts of the next read that has one, and none after the batch's last such read
    (the reference reads past its counts there; zero counts are used here)."""
    colored = [r for r, st in enumerate(stats) if st[0] > 0]
    slots = colored_seq_ids(stats) if quirk2 else None
    lists = []
    for r, (f, nf, inv) in enumerate(stats):
        if first_read_quirk and r == 0:
            f, nf, inv = [(x - 1) % (1 << 64) for x in (f, nf, inv)]
        seq = (f + (nf if inc_nf else 0) + (inv if inc_inv else 0)) % (1 << 64)
        mn = O.lib().orc_color_min_hits(f, nf, inv, tau, int(inc_nf), int(inc_inv))
        assert seq >= 0
        row = r
        if quirk2:
            q = slots[r]
            row = colored[q] if q < len(colored) else None
        if row is None:
            lists.append([])
            continue
        lists.append([c for c in range(counts.shape[1]) if mn > 0 and counts[row, c] >= mn])
    return lists
