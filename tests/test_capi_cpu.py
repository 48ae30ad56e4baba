"""CPU checks of the C ABI library: it loads, exports every declared entry point, and its
host-only functions behave (no GPU compute is called here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "ithaca-sem_amd", "lib", "libithaca_mf.so")
COLL = os.path.join(REPO, "ithaca-sem_amd", "lib", "libithaca_collections.so")


def declared(header):
    src = open(os.path.join(REPO, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:imf|irom)_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header", sorted(h for h in os.listdir(os.path.join(REPO, "include"))
                                          if h.endswith(".h")))
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIB)
    names = declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_collections_library_loads():
    ctypes.CDLL(LIB)
    ctypes.CDLL(COLL)


def test_basis_helper_matches_oracle():
    lib = ctypes.CDLL(LIB)
    for nm, nq in [(2, 2), (5, 6), (6, 7), (8, 16)]:
        arrs = [np.zeros(nq), np.zeros(nq), np.zeros(nq * nq), np.zeros(nm * nq), np.zeros(nm * nq)]
        st = lib.imf_basis_modified_a(nm, nq, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
        assert st == 0
        b = O.basis(nm, nq)
        for a, k in zip(arrs, ["z", "w", "D", "bdata", "dbdata"]):
            scale = np.abs(b["D"]).max()  # D.B sums cancel: error ~ eps * max|D|
            assert np.allclose(a, b[k], atol=1e-14 * scale, rtol=1e-13), k


def test_modified_b_helper_matches_oracle():
    lib = ctypes.CDLL(LIB)
    for nm, nq in [(2, 1), (2, 2), (6, 6), (4, 7), (8, 15)]:
        M = nm * (nm + 1) // 2
        arrs = [np.zeros(nq), np.zeros(nq), np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)]
        st = lib.imf_basis_modified_b(nm, nq, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
        assert st == 0
        b = O.basis_tri(nm, nq + 1, nq)
        for a, k in zip(arrs, ["z1", "w1_raw", "D1", "bdata1", "dbdata1"]):
            scale = max(1.0, np.abs(b["D1"]).max())
            assert np.allclose(a, b[k], atol=1e-14 * scale, rtol=1e-13), (nm, nq, k)


def test_modified_c_helper_matches_oracle():
    lib = ctypes.CDLL(LIB)
    for nm, nq in [(2, 1), (2, 2), (6, 6), (4, 7), (8, 15)]:
        M = nm * (nm + 1) * (nm + 2) // 6
        arrs = [np.zeros(nq), np.zeros(nq), np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)]
        st = lib.imf_basis_modified_c(nm, nq, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
        assert st == 0
        b = O.basis_tet(nm, nq + 1, nq, nq)
        for a, k in zip(arrs, ["z2", "w2_raw", "D2", "bdata2", "dbdata2"]):
            scale = max(1.0, np.abs(b["D2"]).max())
            assert np.allclose(a, b[k], atol=1e-14 * scale, rtol=1e-13), (nm, nq, k)


def test_modified_pyr_c_helper_matches_oracle():
    lib = ctypes.CDLL(LIB)
    for nm, nq in [(2, 1), (2, 2), (6, 6), (4, 7), (8, 15)]:
        M = nm * (nm + 1) * (2 * nm + 1) // 6
        arrs = [np.zeros(nq), np.zeros(nq), np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)]
        st = lib.imf_basis_modified_pyr_c(nm, nq, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
        assert st == 0
        b = O.basis_pyr(nm, nq + 1, nq)
        for a, k in zip(arrs, ["z2", "w2_raw", "D2", "bdata2", "dbdata2"]):
            scale = max(1.0, np.abs(b["D2"]).max())
            assert np.allclose(a, b[k], atol=1e-14 * scale, rtol=1e-13), (nm, nq, k)


def test_create_rejects_bad_requests_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.imf_last_error.restype = ctypes.c_char_p
    nm = (ctypes.c_int * 3)(6, 6, 6)
    nq = (ctypes.c_int * 3)(7, 7, 7)
    nm9 = (ctypes.c_int * 3)(9, 9, 9)
    b = np.zeros(42)
    dp = ctypes.POINTER(ctypes.c_double)
    arr = (dp * 3)(*([b.ctypes.data_as(dp)] * 3))
    h = ctypes.c_void_p()
    st = lib.imf_create(b"Helmholtz_Cube_Deformed", 3, nm, nq, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 2 and b"No such operator key" in lib.imf_last_error()
    st = lib.imf_create(b"Helmholtz_Tet_Deformed", 3, nm9, nq, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 2 and b"combo not implemented" in lib.imf_last_error()
    st = lib.imf_create(b"Helmholtz_Prism_Deformed", 3, nm9, nq, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 2 and b"combo not implemented" in lib.imf_last_error()
    st = lib.imf_create(b"Helmholtz_Prism_Deformed", 2, nm, nq, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 1 and b"dim does not match" in lib.imf_last_error()
    nq13 = (ctypes.c_int * 3)(13, 13, 13)
    st = lib.imf_create(b"Helmholtz_Hex_Deformed", 3, nm, nq13, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 2 and b"combo not initialised" in lib.imf_last_error()
    st = lib.imf_create(b"Helmholtz_Hex_Deformed", 2, nm, nq, arr, arr, arr, arr, 8, 0,
                        ctypes.byref(h))
    assert st == 1
