"""The colors stage's file path on the MI355X through the C ABI: index-file
bodies parsed on the device, color id lists printed on the device, and the
fused per-batch call, bit-exact against the oracle."""
import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
from colorlib import ColoredIndex, kmer_ends_of, read_stats, threshold_lists
from test_oracle_colors import _ascii_body, _binary_body, _random_results

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def fuzz_body(seed, n):
    """Random bytes over the chars that move the ascii automaton."""
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"0123456789--  \n\n\n12a\x80\x00\t", np.uint8)
    return alphabet[rng.integers(0, len(alphabet), n)].tobytes()


@pytest.fixture(scope="module")
def parser():
    return S.IndexFileParser(1 << 20, 1 << 18)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 4095, 4096, 4097, 70_001, 600_000])
def test_parser_fuzz_matches_oracle(parser, n):
    """Segment (16 B) and block (4 KiB) boundaries cut through every kind of token."""
    say(f"parse fuzz n={n}")
    body = fuzz_body(n, n)
    res, ends = parser.parse(0, body)
    r0, e0 = O.parse_index_body(body, 0)
    assert np.array_equal(res, r0) and np.array_equal(ends, e0)


@pytest.mark.parametrize("seed", [1, 2])
def test_parser_printed_files_match_oracle(parser, seed):
    say(f"parse printed seed={seed}")
    res, ke = _random_results(seed, 2000)
    for fmt, body in ((0, _ascii_body(res, ke)), (1, _binary_body(res, ke))):
        r, e = parser.parse(fmt, body)
        assert np.array_equal(r, res) and np.array_equal(e, ke), fmt


@pytest.mark.parametrize("fmt", [0, 1])
def test_batches_cut_at_read_ends_concatenate(parser, fmt):
    """Parsing a body in batch_cut pieces and concatenating equals one parse."""
    say(f"batch cuts fmt={fmt}")
    res, ke = _random_results(7, 3000)
    body = _ascii_body(res, ke) if fmt == 0 else _binary_body(res, ke)
    if fmt == 0:
        body = body.replace(b"\n", b"\n-\n", 5)  # '\n' after '-' is no read end
        res, ke = O.parse_index_body(body, 0)
    got_r, got_e, pos = [], [], 0
    while pos < len(body):
        cut = S.index_file_batch_cut(fmt, body[pos:pos + 5000], 97)
        assert cut > 0
        r, e = parser.parse(fmt, body[pos:pos + cut])
        assert len(e) <= 97
        base = sum(len(x) for x in got_r)
        got_e.append(e + np.uint64(base))
        got_r.append(r)
        pos += cut
    assert np.array_equal(np.concatenate(got_r), res)
    assert np.array_equal(np.concatenate(got_e), ke)


def test_parser_rejects_more_reads_than_its_max():
    say("parser max reads")
    p = S.IndexFileParser(4096, 3)
    with pytest.raises(RuntimeError, match="max reads"):
        p.parse(0, b"1\n2\n3\n4\n")
    r, e = p.parse(0, b"1\n2\n3\n")
    assert list(e) == [1, 2, 3]


@pytest.mark.parametrize("nc", [1, 7, 64, 1000])
def test_color_formatter_matches_oracle(nc):
    say(f"color formatter nc={nc}")
    rng = np.random.default_rng(nc)
    n = 3000
    lens = rng.integers(0, min(nc, 40) + 1, n)
    lens[rng.random(n) < 0.3] = 0
    lists = [np.sort(rng.choice(nc, int(m), replace=False)).astype(np.uint32) for m in lens]
    offsets = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
    ids = np.concatenate(lists) if offsets[-1] else np.zeros(0, np.uint32)
    f = S.ColorFormatter(max(len(ids), 1), n)
    for mode in (0, 1, 2):
        got = f.format(mode, offsets, ids, nc)
        assert got == O.color_lists_format(offsets, ids, nc, mode), mode


@pytest.fixture(scope="module")
def colored(tmp_path_factory):
    ci = ColoredIndex(tmp_path_factory.mktemp("cio"), n_genomes=12, length=15_000, d=6, sub=0.03)
    chars, ends = ci.reads(1200, np.random.default_rng(3).integers(10, 200, 1200), seed=4)
    res = ci.index_results(chars, ends)
    ke = kmer_ends_of(ends)
    return ci, res, ke


@pytest.mark.parametrize("in_fmt", [0, 1])
@pytest.mark.parametrize("tau,inc_nf,inc_inv,quirk", [(0.7, False, False, True),
                                                      (0.3, True, True, False),
                                                      (1.0, False, True, False)])
def test_color_file_batch_matches_oracle(colored, in_fmt, tau, inc_nf, inc_inv, quirk):
    """Index-file body -> parse -> color search -> printer bytes, in one call."""
    ci, res, ke = colored
    say(f"color file batch fmt={in_fmt} tau={tau}")
    body = _ascii_body(res, ke) if in_fmt == 0 else _binary_body(res, ke)
    lists = threshold_lists(ci.oracle_counts(res, ke), read_stats(res, ke), tau, inc_nf, inc_inv,
                            first_read_quirk=quirk)
    offsets = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
    ids = np.array([c for x in lists for c in x], np.uint32)
    col = S.ColorIndex(ci.colors_path)
    fs = S.ColorFileSearcher(col, len(body) + 16, len(ke))
    p = S.color_params(tau, inc_nf, inc_inv, quirk)
    for mode in (0, 1, 2):
        got, n = fs.batch(in_fmt, body, p, mode)
        assert n == len(ke)
        assert got == O.color_lists_format(offsets, ids, ci.num_colors, mode), mode
