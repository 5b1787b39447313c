"""The C-ABI library loads on the CPU host and exports every declared symbol."""
import ctypes
import os
import subprocess

import pytest

import sbwt_amd as S


def test_library_exports_every_header_symbol():
    L = S.lib()
    names = S.header_symbols()
    assert "sbwt_index_create" in names and "sbwt_searcher_search" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "--notes", S.LIB_PATH],
                         capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("llvm-readelf unavailable")
    blob = open(S.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_last_error_starts_empty():
    assert S.lib().sbwt_last_error() in (b"",)


def test_load_rejects_wrong_sbwt_variant(tmp_path):
    """Format errors are reported before any device work (SbwtBuilder.cpp:64,67)."""
    import struct
    p = tmp_path / "bad.sbwt"
    p.write_bytes(struct.pack("<Q", 6) + b"nested" + struct.pack("<Q", 4) + b"v0.1")
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_load(os.fsencode(str(p)), None, 0, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == b"Error input is not a plain-matrix SBWT"
    p.write_bytes(struct.pack("<Q", 12) + b"plain-matrix" + struct.pack("<Q", 4) + b"v0.2")
    rc = S.lib().sbwt_index_load(os.fsencode(str(p)), None, 0, ctypes.byref(h))
    assert rc != 0 and S.lib().sbwt_last_error() == b"Error: wrong SBWT version"


def test_load_rejects_wrong_colors_magic(tmp_path):
    from conftest import FIXTURES
    p = tmp_path / "bad.tcolors"
    import struct
    p.write_bytes(struct.pack("<Q", 5) + b"hello")
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_load(os.fsencode(os.path.join(FIXTURES, "search_test_index.sbwt")),
                                 os.fsencode(str(p)), 0, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == \
        b"The colors file has an incorrect format. Expected 'sdsl-hybrid-v4'"


@pytest.mark.parametrize("kp,k", [(17, 31), (9, 8)])
def test_create_rejects_presearch_letters_out_of_range(kp, k):
    """The presearch prefix is at most min(k, 16) letters; checked before device work."""
    words = (ctypes.c_uint64 * 1)(0b1011)
    acgt = (ctypes.POINTER(ctypes.c_uint64) * 4)(*([ctypes.cast(words, ctypes.POINTER(ctypes.c_uint64))] * 4))
    h = ctypes.c_void_p()
    rc = S.lib().sbwt_index_create(acgt, 4, k, None, None, kp, ctypes.byref(h))
    assert rc != 0
    assert S.lib().sbwt_last_error() == b"presearch letters must be in [1, min(k, 16)]"
