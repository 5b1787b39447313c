"""The sbwt_search command line on the MI355X: byte-identical to the reference's
expected files and to the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_py as O
from conftest import FIXTURES, REPO

pytestmark = pytest.mark.gpu
CLI = os.path.join(REPO, "sbwt-search_amd", "host", "sbwt_search")
NAMES = ["fasta1.fna", "fasta2.fna", "fasta3.fna", "fasta4.fna", "fastq1.fnq"]


def run(args):
    print("[case] sbwt_search " + " ".join(args), flush=True)
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("search", [[], ["--search-algorithm", "streaming"],
                                    ["--search-algorithm", "streaming", "--kmer-table"],
                                    ["--search-algorithm", "adaptive", "--kmer-table"],
                                    ["--search-algorithm", "per-kmer"]],
                         ids=["auto", "streaming", "streaming-table", "adaptive-table", "per-kmer"])
@pytest.mark.parametrize("batch", [None, 40, 31])
def test_index_cli_matches_reference_expected(tmp_path, batch, search):
    sbwt = os.path.join(FIXTURES, "search_test_index.sbwt")
    for name in NAMES:
        out = str(tmp_path / name)
        extra = ["--batch-chars", str(batch)] if batch else []
        run(["index", "-q", os.path.join(FIXTURES, "index_search", name), "-i", sbwt, "-o", out]
            + extra + search)
        exp = open(os.path.join(FIXTURES, "index_search", "expected",
                                name.split(".")[0] + ".txt"), "rb").read()
        assert open(out + ".txt", "rb").read() == exp, name


@pytest.mark.parametrize("streams,repeat", [("3", 1), ("1", 2)])
def test_index_cli_list_streams_and_formats(tmp_path, streams, repeat):
    """Files of a .list over several streams, or many files on one stream (its
    batch buffers are recycled across files); every format equals the oracle's
    files."""
    sbwt = os.path.join(FIXTURES, "search_test_index.sbwt")
    orc = O.OracleSbwt(sbwt)
    names = NAMES * repeat
    ins = [os.path.join(FIXTURES, "index_search", n) for n in names]
    outs = [str(tmp_path / f"{i}_{n}") for i, n in enumerate(names)]
    (tmp_path / "in.list").write_text("\n".join(ins) + "\n")
    (tmp_path / "out.list").write_text("\n".join(outs) + "\n")
    for mode, ext, fmt in (("binary", ".bin", 1), ("bool", ".bool", 2), ("ascii", ".txt", 0)):
        run(["index", "-q", str(tmp_path / "in.list"), "-i", sbwt, "-o", str(tmp_path / "out.list"),
             "-p", mode, "-s", streams, "--batch-chars", "64"])
        for i, o in zip(ins, outs):
            orc.index_file(i, o + ".orc", fmt, True)
            assert open(o + ext, "rb").read() == open(o + ".orc", "rb").read(), (mode, i)


def test_index_cli_gzip_and_no_headers(tmp_path):
    import gzip
    sbwt = os.path.join(FIXTURES, "search_test_index.sbwt")
    src = os.path.join(FIXTURES, "index_search", "fasta2.fna")
    gz = str(tmp_path / "f.fna.gz")
    with open(src, "rb") as f, gzip.open(gz, "wb") as g:
        g.write(f.read())
    run(["index", "-q", gz, "-i", sbwt, "-o", str(tmp_path / "o"), "--no-headers"])
    exp = open(os.path.join(FIXTURES, "index_search", "expected", "fasta2.txt"), "rb").read()
    assert open(str(tmp_path / "o.txt"), "rb").read() == exp[25:]


def test_colors_cli_matches_oracle(tmp_path):
    from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists
    ci = ColoredIndex(tmp_path, n_genomes=6, length=15_000, d=4)
    chars, ends = ci.reads(300, 120, seed=9)
    fa = tmp_path / "r.fna"
    st = 0
    with open(fa, "wb") as f:
        for i, e in enumerate(ends):
            f.write(b">r%d\n" % i + chars[st:int(e)] + b"\n")
            st = int(e)
    run(["index", "-q", str(fa), "-i", ci.sbwt_path, "-k", ci.colors_path, "-o",
         str(tmp_path / "idx")])
    res = ci.index_results(chars, ends)
    assert open(str(tmp_path / "idx.txt"), "rb").read() == \
        O.ascii_index_output(res, ends, 31)
    run(["colors", "-q", str(tmp_path / "idx.txt"), "-k", ci.colors_path, "-o",
         str(tmp_path / "col"), "-t", "0.7"])
    ke = kmer_ends_of(ends)
    exp = threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke), 0.7,
                          first_read_quirk=True)
    body = open(str(tmp_path / "col.txt"), "rb").read()[25:].decode().split("\n")[:-1]
    assert [[int(x) for x in line.split()] for line in body] == exp


@pytest.mark.parametrize("in_mode", ["ascii", "binary"])
def test_colors_cli_formats_batches_and_streams(tmp_path, in_mode):
    """colors over a .list of index files (ascii or binary) in every output format,
    on one stream with small batches (reads of a file spread over many batches) and
    on one stream per file; every output file equals the oracle's."""
    import struct
    from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists
    ci = ColoredIndex(tmp_path, n_genomes=6, length=15_000, d=4)
    ext_in = ".txt" if in_mode == "ascii" else ".bin"
    files = []
    for i in range(3):
        chars, ends = ci.reads(150 + 60 * i, 120, seed=20 + i)
        fa = tmp_path / f"r{i}.fna"
        st = 0
        with open(fa, "wb") as f:
            for j, e in enumerate(ends):
                f.write(b">r%d\n" % j + chars[st:int(e)] + b"\n")
                st = int(e)
        run(["index", "-q", str(fa), "-i", ci.sbwt_path, "-k", ci.colors_path, "-o",
             str(tmp_path / f"idx{i}"), "-p", in_mode])
        res = ci.index_results(chars, ends)
        ke = kmer_ends_of(ends)
        files.append((str(tmp_path / f"idx{i}") + ext_in, res, ke))
    (tmp_path / "in.list").write_text("\n".join(f for f, _, _ in files) + "\n")
    outs = [str(tmp_path / f"col{i}") for i in range(3)]
    (tmp_path / "out.list").write_text("\n".join(outs) + "\n")
    nc = ci.num_colors
    # one stream takes its files largest first (the size-balanced partition)
    first_on_one_stream = max(range(3), key=lambda i: os.path.getsize(files[i][0]))
    for streams, extra in (("1", ["--batch-bytes", "3000"]), ("3", [])):
        for mode, ext, m in (("ascii", ".txt", 0), ("binary", ".bin", 1), ("csv", ".csv", 2)):
            run(["colors", "-q", str(tmp_path / "in.list"), "-k", ci.colors_path, "-o",
                 str(tmp_path / "out.list"), "-p", mode, "-t", "0.7", "-s", streams] + extra)
            for i, (_, res, ke) in enumerate(files):
                # the carried u64 max reaches the first read of each stream only
                quirk = i == first_on_one_stream or streams == "3"
                lists = threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke), 0.7,
                                        first_read_quirk=quirk)
                offsets = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
                ids = np.array([c for x in lists for c in x], np.uint32)
                if m == 2:
                    head = (",".join(str(c) for c in range(1, nc + 1)) + "\n").encode()
                else:
                    name = b"ascii" if m == 0 else b"binary"
                    head = struct.pack("<Q", len(name)) + name + struct.pack("<Q", 4) + b"v1.0"
                want = head + O.color_lists_format(offsets, ids, nc, m)
                assert open(outs[i] + ext, "rb").read() == want, (streams, mode, i)


def test_unreadable_inputs_are_skipped(tmp_path):
    """A missing input file (or an index file of unknown format) is reported and
    skipped; its output file holds the headers only and the other files are
    processed (ContinuousSequenceFileParser.cpp:113-132,
    ContinuousIndexFileParser.cpp:64-100)."""
    import struct
    sbwt = os.path.join(FIXTURES, "search_test_index.sbwt")
    good = os.path.join(FIXTURES, "index_search", "fasta2.fna")
    missing = str(tmp_path / "missing.fna")
    (tmp_path / "in.list").write_text(f"{good}\n{missing}\n{good}\n")
    outs = [str(tmp_path / f"o{i}") for i in range(3)]
    (tmp_path / "out.list").write_text("\n".join(outs) + "\n")
    r = run(["index", "-q", str(tmp_path / "in.list"), "-i", sbwt, "-o",
             str(tmp_path / "out.list"), "-s", "1"])
    assert "missing.fna cannot be opened" in r.stderr
    exp = open(os.path.join(FIXTURES, "index_search", "expected", "fasta2.txt"), "rb").read()
    head = struct.pack("<Q", 5) + b"ascii" + struct.pack("<Q", 4) + b"v1.0"
    assert [open(o + ".txt", "rb").read() for o in outs] == [exp, head, exp]
    # colors: a missing index file and one of unknown format next to a good one
    from colorlib import ColoredIndex
    ci = ColoredIndex(tmp_path, n_genomes=4, length=8_000, d=3)
    chars, ends = ci.reads(40, 100, seed=3)
    fa = tmp_path / "r.fna"
    st = 0
    with open(fa, "wb") as f:
        for j, e in enumerate(ends):
            f.write(b">r%d\n" % j + chars[st:int(e)] + b"\n")
            st = int(e)
    run(["index", "-q", str(fa), "-i", ci.sbwt_path, "-k", ci.colors_path, "-o",
         str(tmp_path / "idx")])
    bad = tmp_path / "bad.txt"
    bad.write_bytes(struct.pack("<Q", 4) + b"text" + struct.pack("<Q", 4) + b"v1.0" + b"1 2\n")
    ins = [str(tmp_path / "idx.txt"), str(tmp_path / "nope.txt"), str(bad)]
    (tmp_path / "cin.list").write_text("\n".join(ins) + "\n")
    couts = [str(tmp_path / f"c{i}") for i in range(3)]
    (tmp_path / "cout.list").write_text("\n".join(couts) + "\n")
    r = run(["colors", "-q", str(tmp_path / "cin.list"), "-k", ci.colors_path, "-o",
             str(tmp_path / "cout.list"), "-s", "1", "-t", "0.7"])
    assert "nope.txt cannot be opened" in r.stderr
    assert "Invalid file format in file: " + str(bad) in r.stderr
    col = [open(o + ".txt", "rb").read() for o in couts]
    assert col[1] == col[2] == head
    assert len(col[0]) > len(head) and col[0].count(b"\n") == len(ends)


@pytest.fixture(scope="module")
def mixed_inputs(tmp_path_factory):
    """A built index and sequence files that take every path of the index
    command's reader: wrapped FASTA (device-parsed chunks), FASTA with a '+'
    line start late in the file and with a record longer than a batch (the
    host reader takes over mid-file), four-line FASTQ (device-parsed), FASTQ
    with records the host reader reads differently (it takes over there),
    CRLF FASTQ without a last '\n', junk before the first header, an empty
    file, and gzip."""
    import gzip
    import workloads as W
    from test_gpu_fasta import fastq_records
    d = tmp_path_factory.mktemp("mixed")
    g = W.genome_random(31, 40_000)
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, 31)
    sbwt = str(d / "m.sbwt")
    b.write(sbwt)
    lens = W.uniform_lengths(8, 1500, 0, 500)
    chars, rends, _ = W.reads(12, g, np.array([len(g)], np.uint64), lens, 0.3, 0.02, 0.01)
    seqs, st = [], 0
    for e in rends:
        seqs.append(chars[st:int(e)].tobytes())
        st = int(e)

    def fasta(recs, width=70):
        out = bytearray()
        for i, s in enumerate(recs):
            out += b">r%d x\n" % i
            for j in range(0, len(s), width):
                out += s[j:j + width] + b"\n"
        return bytes(out)

    files = {
        "wrapped.fna": fasta(seqs),
        "oneline.fna": fasta(seqs[:700], 1 << 20),
        "late_plus.fna": fasta(seqs[:900]) + b"@q\nACGTACGTTT\n+\nIIIIIIIIII\n" + fasta(seqs[900:1000]),
        "long_record.fna": fasta(seqs[:300]) + b">long\n" + g[:9000].tobytes() + b"\n" + fasta(seqs[300:400]),
        "reads.fnq": b"".join(b"@r%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)) for i, s in enumerate(seqs[:500])),
        "junk_first.fna": b"junk line\nmore>" + fasta(seqs[:50])[1:],
        "empty.fna": b"",
        "odd_records.fnq": fastq_records(5, 3000, 0.01),
        "crlf_no_eol.fnq": b"".join(b"@r%d\r\n%s\r\n+\r\n%s\r\n" % (i, s, b"I" * len(s))
                                    for i, s in enumerate(seqs[:300]))[:-1],
    }
    paths = []
    for name, data in files.items():
        p = d / name
        p.write_bytes(data)
        paths.append(str(p))
    gz = d / "wrapped.fna.gz"
    with gzip.open(str(gz), "wb") as f:
        f.write(files["wrapped.fna"])
    paths.append(str(gz))
    return sbwt, paths


@pytest.mark.parametrize("batch", [None, 6000, 700])
@pytest.mark.parametrize("host_parse", [False, True])
def test_index_cli_device_and_host_parsing_match_oracle(tmp_path, mixed_inputs, batch, host_parse):
    sbwt, ins = mixed_inputs
    orc = O.OracleSbwt(sbwt)
    outs = [str(tmp_path / f"o{i}") for i in range(len(ins))]
    (tmp_path / "in.list").write_text("\n".join(ins) + "\n")
    (tmp_path / "out.list").write_text("\n".join(outs) + "\n")
    extra = (["--batch-chars", str(batch)] if batch else []) + (["--host-parse"] if host_parse else [])
    run(["index", "-q", str(tmp_path / "in.list"), "-i", sbwt, "-o", str(tmp_path / "out.list"),
         "-s", "2"] + extra)
    for i, o in zip(ins, outs):
        src = i
        if i.endswith(".gz"):
            import gzip
            src = o + ".plain"
            with gzip.open(i, "rb") as f, open(src, "wb") as g:
                g.write(f.read())
        orc.index_file(src, o + ".orc", 0, True)
        assert open(o + ".txt", "rb").read() == open(o + ".orc", "rb").read(), i
