"""Device formatting of index results into the reference's output bytes."""
import os
import struct

import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-format] {msg}", flush=True)


@pytest.fixture(scope="module")
def small_index(tmp_path_factory):
    d = tmp_path_factory.mktemp("fmt")
    g = W.genome_random(21, 50_000)
    codes, ends = W.concat([g])
    b = W.BuiltSbwt(codes, ends, 31)
    path = str(d / "f.sbwt")
    b.write(path)
    return g, path, O.OracleSbwt(path)


HEADER = {S.FORMAT_ASCII: 5 + 8 + 4 + 8, S.FORMAT_BINARY: 6 + 8 + 4 + 8,
          S.FORMAT_BOOL: 4 + 8 + 4 + 8}


@pytest.mark.parametrize("mode", [S.FORMAT_ASCII, S.FORMAT_BINARY, S.FORMAT_BOOL],
                         ids=["ascii", "binary", "bool"])
def test_formatter_matches_oracle_files(small_index, mode, tmp_path):
    """Reads with not-found and invalid k-mers and reads shorter than k: the
    formatted bytes equal the oracle writer's file after its header."""
    g, path, orc = small_index
    say(f"formatter vs oracle file mode={mode}")
    lens = np.concatenate([np.array([10, 31, 30, 0], np.uint32), W.uniform_lengths(5, 400, 1, 300)])
    chars, ends, _ = W.reads(9, g, np.array([len(g)], np.uint64), lens, 0.3, 0.02, 0.01)
    fasta = tmp_path / "r.fna"
    start = 0
    with open(fasta, "wb") as f:
        for i, e in enumerate(ends):
            f.write(b">r%d\n" % i + chars[start:int(e)].tobytes() + b"\n")
            start = int(e)
    out = tmp_path / "o"
    orc.index_file(str(fasta), str(out), mode, True)
    exp = open(str(out), "rb").read()[HEADER[mode]:]
    idx = S.Index.load(path)
    s = idx.searcher(max(len(chars), 1), len(ends))
    res = s.search_reads(chars.tobytes(), ends)
    lens64 = np.diff(np.concatenate([[0], ends.astype(np.int64)]))
    kends = np.cumsum(np.maximum(lens64 - 30, 0))
    f = S.Formatter(max(len(res), 1), len(ends))
    got = f.format(mode, res, kends, np.ones(len(ends), np.uint8))
    assert got == exp


def test_formatter_ascii_values_and_pieces():
    """Values above 2^32 (64-bit indexes, the walk-stuck sentinel), not-found and
    invalid sentinels, empty pieces and pieces that continue their line."""
    say("formatter ascii edge cases")
    rng = np.random.default_rng(4)
    n_pieces = 3000
    nk = rng.integers(0, 12, n_pieces)
    closes = (rng.random(n_pieces) < 0.6).astype(np.uint8)
    closes[-1] = 1
    res = rng.integers(0, 1 << 63, int(nk.sum()), dtype=np.uint64)
    pick = rng.random(len(res))
    res[pick < 0.2] = O.NOT_FOUND
    res[(pick >= 0.2) & (pick < 0.3)] = O.INVALID
    res[(pick >= 0.3) & (pick < 0.35)] = rng.integers(0, 1 << 32, int(((pick >= 0.3) & (pick < 0.35)).sum()))
    res[(pick >= 0.35) & (pick < 0.37)] = np.uint64(2**64 - 4)
    res[(pick >= 0.37) & (pick < 0.4)] = np.arange(int(((pick >= 0.37) & (pick < 0.4)).sum()))
    kends = np.cumsum(nk)
    for line_open in (False, True):
        # expected: a line per closing piece, its pieces' results joined by ' ';
        # with line_open the first line continues an earlier one
        per_line, cur = [], 0
        starts = np.concatenate([[0], kends[:-1]])
        for e in range(n_pieces):
            cur += int(nk[e])
            if closes[e]:
                per_line.append(cur)
                cur = 0
        line_ends = np.cumsum(per_line)
        exp = O.ascii_index_output(res, line_ends, 1, headers=False)
        if line_open and per_line[0] > 0:
            exp = b" " + exp
        f = S.Formatter(len(res), n_pieces)
        got = f.format(S.FORMAT_ASCII, res, kends, closes, line_open=line_open)
        assert got == exp, (line_open, got[:200], exp[:200])
        del starts


def test_formatter_sizes_before_writing():
    f = S.Formatter(16, 4)
    import torch
    res = torch.tensor([5, 123456], dtype=torch.int64, device="cuda")
    ke = torch.tensor([2], dtype=torch.int64, device="cuda")
    cl = torch.tensor([1], dtype=torch.uint8, device="cuda")
    n = f.format_device(S.FORMAT_ASCII, res.data_ptr(), ke.data_ptr(), cl.data_ptr(), 1, False)
    assert n == len(b"5 123456\n")
    out = torch.full((n + 4,), 7, dtype=torch.uint8, device="cuda")
    assert f.format_device(S.FORMAT_ASCII, res.data_ptr(), ke.data_ptr(), cl.data_ptr(), 1, False,
                           out.data_ptr(), n - 1) == n
    assert (out.cpu().numpy() == 7).all()  # too small: nothing written
    f.format_device(S.FORMAT_ASCII, res.data_ptr(), ke.data_ptr(), cl.data_ptr(), 1, False,
                    out.data_ptr(), n)
    assert out.cpu().numpy().tobytes()[:n] == b"5 123456\n"
