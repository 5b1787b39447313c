"""FASTA chunks parsed on the MI355X through the C ABI: chars and read ends
equal the oracle's record reader, chunks cut by sbwt_fasta_chunk_cut
concatenate to the whole file, and the fused parse -> search -> format call
writes the oracle's index output bytes."""
import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W

pytestmark = pytest.mark.gpu

HEADER = {S.FORMAT_ASCII: 5 + 8 + 4 + 8, S.FORMAT_BINARY: 6 + 8 + 4 + 8,
          S.FORMAT_BOOL: 4 + 8 + 4 + 8}


def say(msg):
    print(f"[gpu-fasta] {msg}", flush=True)


def fuzz_fasta(seed, n):
    """Random FASTA-like bytes: headers ('>' and '@' at line starts), empty
    lines, '\\r', other bytes; no line starts with '+'; starts with '>'."""
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGTACGTACGTNacgt>@\n\n\n\r x+\x80\x00", np.uint8)
    b = alphabet[rng.integers(0, len(alphabet), n)]
    if n:
        b[0] = ord(">")
    nl = np.flatnonzero(b[:-1] == ord("\n")) + 1
    b[nl[b[nl] == ord("+")]] = ord("A")
    return b.tobytes()


def oracle_parse(tmp_path, raw):
    p = tmp_path / "f.fna"
    p.write_bytes(raw)
    chars, ends = O.read_seqfile(str(p))
    return bytes(chars), np.asarray(ends, np.uint64)


@pytest.fixture(scope="module")
def parser():
    return S.FastaParser(1 << 21, 1 << 18)


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 4095, 4096, 4097, 70_001, 1_500_000])
def test_parser_fuzz_matches_oracle(parser, tmp_path, n):
    """Segment (16 B) and block (4 KiB) boundaries cut through headers, empty
    lines and sequence lines."""
    say(f"fasta fuzz n={n}")
    raw = fuzz_fasta(n, n)
    chars, ends = parser.parse(raw)
    c0, e0 = oracle_parse(tmp_path, raw)
    assert chars == c0 and np.array_equal(ends, e0)


def test_parser_edge_records(parser, tmp_path):
    say("fasta edge records")
    for raw in [b">", b">h", b">h\n", b">h\nACGT", b">h\nACGT\n", b">\n>\n>\n", b"@a\nAC\n>b\n\nGT\n",
                b">h\r\nAC\r\nGT\r\n", b">h\nAC>GT\n@x\nA@C\n", b">a\n\n\n\n>b\nA\n" * 50]:
        chars, ends = parser.parse(raw)
        c0, e0 = oracle_parse(tmp_path, raw)
        assert chars == c0 and np.array_equal(ends, e0), raw


def test_parser_rejects_what_the_grammar_leaves_to_the_host(tmp_path):
    say("fasta rejections")
    p = S.FastaParser(4096, 3)
    with pytest.raises(RuntimeError, match="header"):
        p.parse(b"ACGT\n>h\nAC\n")
    with pytest.raises(RuntimeError, match="'\\+'"):
        p.parse(b"@r\nACGT\n+\nIIII\n")
    with pytest.raises(RuntimeError, match="max reads"):
        p.parse(b">1\nA\n>2\nA\n>3\nA\n>4\nA\n")
    chars, ends = p.parse(b">1\nA\n>2\nA\n>3\nA\n")
    assert chars == b"AAA" and list(ends) == [1, 2, 3]
    chars, ends = p.parse(b"")
    assert chars == b"" and len(ends) == 0


@pytest.mark.parametrize("max_reads,window", [(1 << 18, 5000), (7, 100_000), (1, 3000)])
def test_chunks_cut_at_records_concatenate(parser, tmp_path, max_reads, window):
    """Parsing a file in chunk_cut pieces and concatenating equals one parse."""
    say(f"fasta chunk cuts max_reads={max_reads} window={window}")
    raw = fuzz_fasta(99, 200_000)
    c0, e0 = oracle_parse(tmp_path, raw)
    got_c, got_e, pos = [], [], 0
    while pos < len(raw):
        piece = raw[pos:pos + window]
        cut = S.fasta_chunk_cut(piece, pos + window >= len(raw), max_reads)
        assert cut > 0
        c, e = parser.parse(piece[:cut])
        assert len(e) <= max_reads
        base = sum(len(x) for x in got_c)
        got_e.append(e + np.uint64(base))
        got_c.append(c)
        pos += cut
    assert b"".join(got_c) == c0
    assert np.array_equal(np.concatenate(got_e), e0)


@pytest.fixture(scope="module")
def small_index(tmp_path_factory):
    d = tmp_path_factory.mktemp("fasta")
    g = W.genome_random(23, 60_000)
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(d / "f.sbwt")
    b.write(path)
    return g, path, O.OracleSbwt(path)


def write_fasta(path, chars, ends, width):
    """Records with headers, sequence wrapped at `width` chars per line."""
    start = 0
    with open(path, "wb") as f:
        for i, e in enumerate(ends):
            s = chars[start:int(e)].tobytes()
            f.write(b">r%d desc\n" % i)
            for j in range(0, len(s), width):
                f.write(s[j:j + width] + b"\n")
            start = int(e)


@pytest.mark.parametrize("mode", [S.FORMAT_ASCII, S.FORMAT_BINARY, S.FORMAT_BOOL],
                         ids=["ascii", "binary", "bool"])
@pytest.mark.parametrize("width", [60, 1 << 20])
def test_search_fasta_format_matches_oracle_file(small_index, tmp_path, mode, width):
    """Raw FASTA chunk -> device parse -> search -> format: the oracle's index
    output file after its header, including reads shorter than k, empty reads,
    not-found and invalid k-mers, and wrapped sequence lines."""
    g, path, orc = small_index
    say(f"search fasta format mode={mode} width={width}")
    lens = np.concatenate([np.array([10, 31, 30, 0, 32], np.uint32),
                           W.uniform_lengths(6, 3000, 1, 400)])
    chars, ends, _ = W.reads(11, g, np.array([len(g)], np.uint64), lens, 0.3, 0.02, 0.01)
    fasta = tmp_path / "r.fna"
    write_fasta(fasta, chars, ends, width)
    out = tmp_path / "o"
    orc.index_file(str(fasta), str(out), mode, True)
    exp = open(str(out), "rb").read()[HEADER[mode]:]
    raw = fasta.read_bytes()
    idx = S.Index.load(path)
    s = idx.searcher(len(raw), len(ends))
    f = S.Formatter(len(raw), len(ends))
    p = S.FastaParser(len(raw), len(ends))
    got, n = p.search_format(s, f, mode, raw)
    assert n == len(ends)
    assert got == exp
    # chunked by chunk_cut: the per-chunk bytes concatenate to the same file
    parts, pos = [], 0
    while pos < len(raw):
        piece = raw[pos:pos + 20_000]
        cut = S.fasta_chunk_cut(piece, pos + 20_000 >= len(raw), len(ends))
        assert cut > 0
        parts.append(p.search_format(s, f, mode, piece[:cut])[0])
        pos += cut
    assert b"".join(parts) == exp


def fastq_records(seed, n_rec, odd=0.0):
    """FASTQ text of n_rec records. With probability `odd` a record takes one
    of the forms the host reader reads differently from four lines (empty
    sequence, short quality, sequence starting with '+', '@' or '>', a
    second sequence line, junk before the header) or a harmless variant
    (CRLF endings, '>' header, '+name' line, quality longer than the
    sequence, '@' / '>' / '+' inside the quality)."""
    rng = np.random.default_rng(seed)
    out = bytearray()
    qual_chars = np.frombuffer(b"IIII#5@>+!?", np.uint8)
    for i in range(n_rec):
        L = int(rng.integers(1, 120))
        seq = bytes(np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, L)])
        qual = bytes(qual_chars[rng.integers(0, len(qual_chars), L)])
        head, plus, nl, extra = b"@r%d" % i, b"+", b"\n", b""
        if rng.random() < odd:
            kind = int(rng.integers(0, 11))
            if kind == 0:
                seq, qual = b"", b"II@II"
            elif kind == 1:
                qual = qual[:-1]
            elif kind == 2:
                seq = b"+@>"[int(rng.integers(0, 3))].to_bytes(1, "little") + seq
                qual += b"I"
            elif kind == 3:
                extra = b"ACGT\n"
                qual += b"IIII"
            elif kind == 4:
                head = b"junk\n" + head
            elif kind == 5:
                nl = b"\r\n"
            elif kind == 6:
                head = b">r%d" % i
            elif kind == 7:
                plus = b"+r%d" % i
            elif kind == 8:
                qual += b"@@>"
            elif kind == 9:
                qual = b"@" + qual[1:]
            else:
                seq, qual = b"", b""
        out += head + nl + seq + nl + extra + plus + nl + qual + nl
    return bytes(out)


def parse_in_chunks(parser, raw, window, tmp_path):
    """The command line's protocol: parse whole-record prefixes chunk by chunk
    and give the rest to the host reader (here the oracle) once a chunk yields
    nothing."""
    chars, ends, pos, device_reads = [], [], 0, 0
    while pos < len(raw):
        piece = raw[pos:pos + window]
        used, c, e = parser.parse(piece, pos + window >= len(raw))
        if used == 0:
            c, e = oracle_parse(tmp_path, raw[pos:])
            pos = len(raw)
        else:
            device_reads += len(e)
            pos += used
        base = sum(len(x) for x in chars)
        ends.append(np.asarray(e, np.uint64) + np.uint64(base))
        chars.append(c)
    return b"".join(chars), np.concatenate(ends) if ends else np.zeros(0, np.uint64), device_reads


@pytest.mark.parametrize("odd,window,max_reads", [(0.0, 1 << 21, 1 << 18), (0.0, 3000, 1 << 18),
                                                  (0.0, 1 << 21, 5), (0.02, 1 << 21, 1 << 18),
                                                  (0.3, 5000, 1 << 18)])
def test_fastq_chunks_match_oracle(tmp_path, odd, window, max_reads):
    """Four-line records parse on the device up to the first record the host
    reader reads differently; the host (oracle) parse of the rest completes
    the file: together they equal the oracle's parse of the whole file."""
    say(f"fastq chunks odd={odd} window={window} max_reads={max_reads}")
    raw = fastq_records(int(odd * 100) + window % 97, 4000, odd)
    p = S.FastqParser(1 << 21, max_reads)
    chars, ends, dev = parse_in_chunks(p, raw, window, tmp_path)
    c0, e0 = oracle_parse(tmp_path, raw)
    assert chars == c0 and np.array_equal(ends, e0)
    if odd == 0.0:
        assert dev == 4000  # every record parsed on the device
    else:
        assert dev > 0


def test_fastq_record_forms(tmp_path):
    """Each form alone: the device takes exactly the records the host reader
    reads as four lines; a last record without '\\n' and trailing empty lines
    are handled as the host reader handles them."""
    say("fastq record forms")
    p = S.FastqParser(1 << 16, 1 << 10)
    good = b"@a\nACGT\n+\nIIII\n"
    cases = [
        (good, 1), (good * 3, 3), (good + b"@b\nAC\n+\nII", 2), (good + b"\n\n", 1),
        (good + b"@b\n\n+\n\n", 1), (good + b"@b\nACG\n+\nII\n" + good, 1),
        (good + b"@b\n+AC\n+\nIII\n", 1), (good + b"@b\nAC\nGT\n+\nIIII\n", 1),
        (good + b"@b\nACGT\n+\nIIIIII@@\n" + good, 3), (b">a\nACGT\n+x\nI@>+\n", 1),
        (b"@a\r\nACGT\r\n+\r\nIIII\r\n", 1), (b"@a\nACGT\n-\nIIII\n", 0),
    ]
    for raw, device_records in cases:
        used, c, e = p.parse(raw, True)
        assert len(e) == device_records, raw
        chars, ends, _ = parse_in_chunks(p, raw, 1 << 16, tmp_path)
        c0, e0 = oracle_parse(tmp_path, raw)
        assert chars == c0 and np.array_equal(ends, e0), raw
    # not the end of the file: a last line without '\n' may continue
    used, c, e = p.parse(good + b"@b\nAC\n+\nII", False)
    assert used == len(good) and len(e) == 1


@pytest.mark.parametrize("mode", [S.FORMAT_ASCII, S.FORMAT_BOOL], ids=["ascii", "bool"])
def test_search_fastq_format_matches_oracle_file(small_index, tmp_path, mode):
    g, path, orc = small_index
    say(f"search fastq format mode={mode}")
    lens = np.concatenate([np.array([10, 31, 30, 1, 32], np.uint32), W.uniform_lengths(7, 3000, 1, 400)])
    chars, ends, _ = W.reads(13, g, np.array([len(g)], np.uint64), lens, 0.3, 0.02, 0.01)
    fq = tmp_path / "r.fq"
    st = 0
    with open(fq, "wb") as f:
        for i, e in enumerate(ends):
            s = chars[st:int(e)].tobytes()
            f.write(b"@r%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)))
            st = int(e)
    out = tmp_path / "o"
    orc.index_file(str(fq), str(out), mode, True)
    exp = open(str(out), "rb").read()[HEADER[mode]:]
    raw = fq.read_bytes()
    idx = S.Index.load(path)
    s = idx.searcher(len(raw), len(ends))
    f = S.Formatter(len(raw), len(ends))
    p = S.FastqParser(len(raw), len(ends))
    got, used, n = p.search_format(s, f, mode, raw)
    assert used == len(raw) and n == len(ends)
    assert got == exp
    parts, pos = [], 0
    while pos < len(raw):
        piece = raw[pos:pos + 30_000]
        b, used, _ = p.search_format(s, f, mode, piece, pos + 30_000 >= len(raw))
        assert used > 0
        parts.append(b)
        pos += used
    assert b"".join(parts) == exp
