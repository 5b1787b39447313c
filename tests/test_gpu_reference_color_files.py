"""The reference's own color-stage inputs (color_search/*.indexes.txt) through
`sbwt_search colors` and the fused colors batch, with --include-not-found, against
the oracle.

The reference's expected color files need its absent themisto index, so the
color index here is built from seeded genomes with more nodes than the largest
index in those files (19,359,334). With --include-not-found a read whose indexes
are all -1 has a positive threshold and is printed with the counts of the next
read that has a found index (quirk 2: colored_seq_id, IndexFileParser.cpp:65-74,
ContinuousColorResultsPrinter.hpp:201-209); the expected lists follow the
oracle's warp layout, which is pinned by the reference parser tests' vectors.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W
from colorlib import colored_seq_ids, read_stats
from conftest import FIXTURES, REPO

pytestmark = pytest.mark.gpu
CLI = os.path.join(REPO, "sbwt-search_amd", "host", "sbwt_search")
NAMES = ["fasta1", "fasta2", "fasta3", "fasta4"]


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def big_colors(tmp_path_factory):
    import ctypes as C
    tmp = tmp_path_factory.mktemp("refcol")
    say("color index of 6 genomes (3 roots x 3.5 Mbp + 2% mutants)")
    gens = []
    for r in range(3):
        root = W.genome_random(900 + r, 3_500_000)
        gens += [root, W.genome_mutate(root, 910 + r, 0.02)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    assert b.n > 19_359_334
    col = b.colors(np.arange(len(gens), dtype=np.uint32), d=20)
    path = str(tmp / "big.tcolors")
    col.write(path)
    oc = O.Colors()
    assert O.lib().orc_colors_load(path.encode(), C.byref(oc)) == 0
    return path, oc


def _expected(oc, body, tau, inc_nf):
    res, kends = O.parse_index_body(body, 0)
    n = len(kends)
    nc = int(oc.num_colors)
    counts = np.zeros(max(n, 1) * nc, np.uint64)
    import ctypes as C
    O.lib().orc_color_counts(C.byref(oc), O.p64(np.ascontiguousarray(res, np.uint64)),
                             O.p64(np.ascontiguousarray(kends, np.uint64)), n, O.p64(counts), 4)
    counts = counts[:n * nc].reshape(n, nc)
    stats = read_stats(res, kends)
    slots = colored_seq_ids(stats)
    colored = [r for r, st in enumerate(stats) if st[0] > 0]
    lists, borrowed = [], 0
    for r, (f, nf, inv) in enumerate(stats):
        if r == 0:  # the reference printer's initial carry (quirk 1)
            f, nf, inv = [(x - 1) % (1 << 64) for x in (f, nf, inv)]
        mn = O.lib().orc_color_min_hits(f, nf, inv, tau, int(inc_nf), 0)
        q = slots[r]
        if mn == 0 or q >= len(colored):
            lists.append([])
            continue
        hit = np.flatnonzero(counts[colored[q]] >= mn).tolist()
        if stats[r][0] == 0 and hit:  # a read with no found index prints colors
            borrowed += 1
        lists.append(hit)
    return lists, borrowed


def _with_blank_reads(data):
    """The file with a read of all -1 before every read and one at the end (the
    reference's files have only empty blank reads, whose threshold is 0)."""
    lines = data[25:].split(b"\n")[:-1]
    out = []
    for i, line in enumerate(lines):
        out.append(b" ".join([b"-1"] * (3 + i % 4)))
        out.append(line)
    out.append(b"-1 -1")
    return data[:25] + b"".join(x + b"\n" for x in out)


@pytest.mark.parametrize("variant", ["reference", "with-blank-reads"])
@pytest.mark.parametrize("inc_nf", [True, False], ids=["include-not-found", "default"])
def test_reference_index_files_through_colors_cli(big_colors, tmp_path, inc_nf, variant):
    path, oc = big_colors
    total_borrowed = 0
    for name in NAMES:
        src = os.path.join(FIXTURES, "color_search", f"{name}.indexes.txt")
        data = open(src, "rb").read()
        if variant != "reference":
            data = _with_blank_reads(data)
            src = str(tmp_path / f"{name}.blank.txt")
            open(src, "wb").write(data)
        lists, borrowed = _expected(oc, data[25:], 0.7, inc_nf)
        total_borrowed += borrowed
        out = str(tmp_path / name)
        args = [CLI, "colors", "-q", src, "-k", path, "-o", out, "-t", "0.7"]
        if inc_nf:
            args.append("--include-not-found")
        say("sbwt_search " + " ".join(args[1:]))
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        body = open(out + ".txt", "rb").read()[25:].decode().split("\n")[:-1]
        assert [[int(x) for x in line.split()] for line in body] == lists, name
        # the fused library call on the same body
        col = S.ColorIndex(path)
        fs = S.ColorFileSearcher(col, len(data) + 16, 64)
        got, n = fs.batch(0, data[25:], S.color_params(0.7, inc_nf, False, True), 0)
        offsets = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
        ids = np.array([c for x in lists for c in x], np.uint32)
        assert got == O.color_lists_format(offsets, ids, int(oc.num_colors), 0), name
    if inc_nf and variant != "reference":
        # reads of all -1 borrowed a later read's counts: quirk 2 was exercised
        assert total_borrowed > 0
    elif not inc_nf:
        assert total_borrowed == 0


def test_blank_reads_stay_with_their_colored_read_across_batches(big_colors, tmp_path):
    """Small --batch-bytes: the colors command ends a batch after a read with a
    found index, so reads of all -1 keep the next colored read in their batch and
    the output equals the whole-file layout."""
    path, oc = big_colors
    src = os.path.join(FIXTURES, "color_search", "fasta2.indexes.txt")
    data = _with_blank_reads(open(src, "rb").read())
    body = data[25:]
    lines = body.split(b"\n")[:-1]
    # the file's reads repeated, so many batches end near blank reads
    rep = b"".join(line + b"\n" for line in lines) * 30
    f = tmp_path / "rep.txt"
    f.write_bytes(data[:25] + rep)
    lists, borrowed = _expected(oc, rep, 0.7, True)
    assert borrowed > 0
    for bb in ("1200", "3000", "100000"):
        out = str(tmp_path / f"rep{bb}")
        args = [CLI, "colors", "-q", str(f), "-k", path, "-o", out, "-t", "0.7",
                "--include-not-found", "--batch-bytes", bb]
        say("sbwt_search " + " ".join(args[1:]))
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = open(out + ".txt", "rb").read()[25:].decode().split("\n")[:-1]
        assert [[int(x) for x in line.split()] for line in got] == lists, bb
