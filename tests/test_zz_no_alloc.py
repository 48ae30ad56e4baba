"""The allocation-free eigensolver entry point (irom_eig_sym_noalloc, static device
workspace for n <= 512) against irom_eig_sym on the same matrices, called through the C ABI
twice in a row (the second call smaller, on the workspace the first one left behind)."""
import ctypes

import numpy as np
import pytest
import torch

from ithaca_hip import pod

pytestmark = pytest.mark.gpu


def _lib():
    L = pod._lib()
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.irom_eig_sym_noalloc.argtypes = [i, vp, vp, vp, d, i, ctypes.POINTER(i), vp]
    L.irom_eig_sym_noalloc.restype = i
    L.irom_eig_sym.argtypes = [i, vp, vp, vp, d, i, ctypes.POINTER(i), vp]
    L.irom_eig_sym.restype = i
    return L


def _psd(n, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.logspace(4, -8, n // 2), 1e-10 * rng.uniform(0, 1, n - n // 2)])
    G = (U * lam) @ U.T
    return torch.from_numpy(0.5 * (G + G.T)).cuda()


def _call(fn, G):
    n = G.shape[0]
    ev = np.zeros(n)
    V = torch.empty((n, n), dtype=torch.float64, device="cuda")
    sw = ctypes.c_int()
    tol = max(1e-14, n * 2.220446049250313e-16)
    st = fn(n, ctypes.c_void_p(G.data_ptr()), ev.ctypes.data_as(ctypes.c_void_p),
            ctypes.c_void_p(V.data_ptr()), tol, 60, ctypes.byref(sw),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, ev, V.cpu().numpy(), sw.value


def test_noalloc_eig_matches_the_allocating_entry_point():
    L = _lib()
    for n, seed in ((512, 1), (200, 2)):
        print(f"[case] irom_eig_sym_noalloc n={n}", flush=True)
        G = _psd(n, seed)
        st0, ev0, V0, sw0 = _call(L.irom_eig_sym, G)
        st1, ev1, V1, sw1 = _call(L.irom_eig_sym_noalloc, G)
        assert st0 == 0 and st1 == 0, L.irom_eig_last_error()
        print(f"sweeps {sw0} / {sw1}, bitwise equal: "
              f"{np.array_equal(ev0, ev1) and np.array_equal(V0, V1)}")
        assert sw1 == sw0
        assert np.abs(ev1 - ev0).max() <= 1e-13 * ev0[0]
        assert np.abs(V1 - V0).max() <= 1e-12
        Gh = G.cpu().numpy()
        # the eigenpairs themselves (V row j is the eigenvector of ev[j])
        assert np.linalg.norm(Gh @ V1.T - V1.T * ev1, axis=0).max() <= 1e-11 * ev1[0]


def test_noalloc_eig_rejects_orders_beyond_its_workspace():
    L = _lib()
    G = _psd(513, 3)
    st, _, _, _ = _call(L.irom_eig_sym_noalloc, G)
    assert st == 1
    assert b"n <= 512" in L.irom_eig_last_error()
