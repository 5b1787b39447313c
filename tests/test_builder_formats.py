"""The .sbwt writer and the host builder against a file the reference ships:
tests/golden/reference_fixtures/search_test_index.sbwt (k = 30, no reverse
complements, 236 nodes) was written by the SBWT tool from
search_test_indexed.fna. Building the same FASTA here and writing it gives the
same bytes: the sdsl bit_vector and rank_support_v5 layouts (which the .tcolors
writer and both loaders share), the suffix-group starts, the C map, the 8-mer
interval table and the trailing fields. The GPU builder's files equal the host
builder's (tests/test_gpu_builder.py), so this pins both.
"""
import os

import numpy as np

import oracle_py as O
import workloads as W
from conftest import FIXTURES


def _codes(chars):
    m = np.full(256, 4, np.uint8)
    for i, c in enumerate(b"ACGT"):
        m[c] = i
        m[c + 32] = i
    return m[np.frombuffer(bytes(chars), np.uint8)]


def test_host_build_of_fixture_fasta_writes_the_reference_file(tmp_path):
    chars, ends = O.read_seqfile(os.path.join(FIXTURES, "search_test_indexed.fna"))
    built = W.BuiltSbwt(_codes(chars), ends, 30, add_rc=False)
    assert (built.n, built.n_kmers) == (236, 177)
    out = tmp_path / "ours.sbwt"
    built.write(str(out))
    ref = open(os.path.join(FIXTURES, "search_test_index.sbwt"), "rb").read()
    assert out.read_bytes() == ref


def test_reference_file_changes_are_seen(tmp_path):
    """A flipped bit in the file's A vector makes the bytes differ (the
    comparison above is not vacuous)."""
    chars, ends = O.read_seqfile(os.path.join(FIXTURES, "search_test_indexed.fna"))
    built = W.BuiltSbwt(_codes(chars), ends, 30, add_rc=False)
    out = tmp_path / "ours.sbwt"
    built.write(str(out))
    ref = bytearray(open(os.path.join(FIXTURES, "search_test_index.sbwt"), "rb").read())
    ref[32 + 8] ^= 1  # first word of the A vector
    assert out.read_bytes() != bytes(ref)
