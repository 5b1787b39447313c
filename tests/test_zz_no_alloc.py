"""The allocation-free fused search (sbwt_index_search_reads_noalloc) through the
C ABI against the searcher's entry point (sbwt_searcher_search_reads_device) and
the oracle: two calls in a row on one GPU, the second with a smaller batch (the
shared scratch keeps the first call's contents), with and without the key-k-mer
move, on a narrow and a 64-bit index; and a batch past the maximum is rejected.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import sbwt_amd as S
import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def colored(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("noalloc")
    root = W.genome_random(71, 60_000)
    gens = [root] + [W.genome_mutate(root, s, 0.02) for s in (72, 73, 74)]
    codes, ends = W.concat(gens)
    b = W.BuiltSbwt(codes, ends, 31)
    col = b.colors(np.arange(len(gens), dtype=np.uint32), d=9)
    sp, cp = str(tmp / "n.sbwt"), str(tmp / "n.tcolors")
    b.write(sp)
    col.write(cp)
    km = np.concatenate([col.key_marks(), np.zeros(8, np.uint64)])
    return codes, ends, sp, cp, O.OracleSbwt(sp), km


def _batch(codes, ends, n, seed):
    lens = W.uniform_lengths(seed, n, 20, 400)
    chars, rends, _ = W.reads(seed, codes, ends, lens, 0.1, 0.01, 0.002)
    return chars.tobytes(), rends


@pytest.mark.parametrize("wide", [False, True], ids=["narrow", "wide"])
@pytest.mark.parametrize("key", [False, True], ids=["plain", "key-move"])
def test_noalloc_equals_searcher_twice(colored, key, wide, monkeypatch):
    import torch
    codes, ends, sp, cp, orc, km = colored
    if wide:
        monkeypatch.setenv("SBWT_AMD_FORCE_WIDE", "1")
    idx = S.Index.load(sp, cp)
    assert idx.wide == wide
    say(f"no-alloc search key={key} wide={wide}")
    for n, seed in ((9000, 81), (700, 82)):  # the second batch is smaller
        chars, rends = _batch(codes, ends, n, seed)
        exp = orc.index_reads(chars, rends, key_marks=km if key else None)
        dc = torch.frombuffer(bytearray(chars), dtype=torch.uint8).cuda()
        de = torch.from_numpy(rends.astype(np.int64)).cuda()
        got = torch.full((len(chars) + 8,), -5, dtype=torch.int64, device="cuda")
        ref = torch.empty(len(chars) + 8, dtype=torch.int64, device="cuda")
        stream = torch.cuda.Stream()
        torch.cuda.synchronize()  # the inputs were written on torch's stream
        idx.search_reads_noalloc(dc.data_ptr(), len(chars), de.data_ptr(), len(rends),
                                 got.data_ptr(), move_to_key_kmer=key, stream=stream.cuda_stream)
        stream.synchronize()
        s = idx.searcher(len(chars), len(rends), move_to_key_kmer=key)
        s.set_algorithm(S.ALGO_STREAMING)
        nk = s.search_reads_device(dc.data_ptr(), len(chars), de.data_ptr(), len(rends),
                                   ref.data_ptr(), want_count=True)
        assert nk == len(exp)
        g = got.cpu().numpy()
        assert np.array_equal(g[:nk].view(np.uint64), ref[:nk].cpu().numpy().view(np.uint64))
        assert np.array_equal(g[:nk].view(np.uint64), exp)
        assert (g[nk:] == -5).all()  # nothing written past the k-mers


def test_noalloc_rejects_batches_past_the_maximum(colored):
    codes, ends, sp, cp, orc, km = colored
    idx = S.Index.load(sp)
    say("no-alloc search: oversized batch rejected")
    L = S.lib()
    rc = L.sbwt_index_search_reads_noalloc(idx.h, 0, C.c_void_p(16), S.NOALLOC_MAX_CHARS + 1,
                                           C.c_void_p(16), 1, C.c_void_p(16), None)
    assert rc != 0 and b"maximum" in L.sbwt_last_error()
    rc = L.sbwt_index_search_reads_noalloc(idx.h, 0, C.c_void_p(16), 100, C.c_void_p(16),
                                           S.NOALLOC_MAX_READS + 1, C.c_void_p(16), None)
    assert rc != 0 and b"maximum" in L.sbwt_last_error()
    rc = L.sbwt_index_search_reads_noalloc(idx.h, 1, C.c_void_p(16), 100, C.c_void_p(16), 1,
                                           C.c_void_p(16), None)
    assert rc != 0 and b"key-k-mer marks" in L.sbwt_last_error()
