"""FP64 CPU oracle of the reduced-basis kernels (TEST INFRASTRUCTURE).

* ``solve``        : A(mu_i) x_i = b_i by LAPACK LU (numpy.linalg.solve), per instance.
* ``solve_eigen``  : the reference's own reduced solver, Eigen 3.3.7
                     ColPivHouseholderQR::solve (CoupledLinearNS_ROM.cpp:4053), through
                     oracle/_ref/libref_eigen.so when that library has been built.
* ``assemble``     : gen_affine_mat_proj (CoupledLinearNS_ROM.cpp:4694-4704) restated:
                     A(mu) = sum_q theta_q A_q.
* ``param_grad``   : dJ/dmu_q = -lambda^T A_q x, with A^T lambda = l (J = l^T x).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "_ref", "libref_eigen.so")
_EIG = None


def eigen_lib():
    """The Eigen reference library, or None when oracle/_ref has not been built."""
    global _EIG
    if _EIG is None and os.path.exists(REF):
        _EIG = ctypes.CDLL(REF)
    return _EIG


def instance(batch, i):
    n, q = int(batch["N"][i]), int(batch["Q"][i])
    a0 = int(batch["a_off"][i])
    Aq = batch["A"][a0:a0 + q * n * n].reshape(q, n, n)
    th = batch["theta"][int(batch["t_off"][i]):int(batch["t_off"][i]) + q]
    v0 = int(batch["v_off"][i])
    return Aq, th, slice(v0, v0 + n)


def assemble(Aq, th):
    A = np.zeros(Aq.shape[1:])
    for q in range(len(th)):
        A += th[q] * Aq[q]
    return A


def solve(batch, rhs=None, transpose=False, idx=None):
    rhs = batch["b"] if rhs is None else rhs
    x = np.zeros_like(rhs)
    for i in (range(len(batch["N"])) if idx is None else idx):
        Aq, th, sl = instance(batch, i)
        A = assemble(Aq, th)
        x[sl] = np.linalg.solve(A.T if transpose else A, rhs[sl])
    return x


def solve_eigen(batch, idx):
    L = eigen_lib()
    if L is None:
        return None
    dp = ctypes.POINTER(ctypes.c_double)
    x = np.zeros_like(batch["b"])
    for i in idx:
        Aq, th, sl = instance(batch, i)
        A = np.ascontiguousarray(assemble(Aq, th))
        bb = np.ascontiguousarray(batch["b"][sl])
        xi = np.zeros(len(bb))
        L.ref_colpivqr_solve(len(bb), A.ctypes.data_as(dp), bb.ctypes.data_as(dp),
                             xi.ctypes.data_as(dp))
        x[sl] = xi
    return x


def eigen_colpivqr(A, b):
    """x = A.colPivHouseholderQr().solve(b) with the reference's Eigen (row-major A)."""
    L = eigen_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    A = np.ascontiguousarray(A)
    b = np.ascontiguousarray(b)
    x = np.zeros(len(b))
    L.ref_colpivqr_solve(len(b), A.ctypes.data_as(dp), b.ctypes.data_as(dp), x.ctypes.data_as(dp))
    return x


def param_grad(batch, lam, x):
    g = np.zeros(int(np.sum(batch["Q"])))
    for i in range(len(batch["N"])):
        Aq, th, sl = instance(batch, i)
        t0 = int(batch["t_off"][i])
        for q in range(len(th)):
            g[t0 + q] = -lam[sl] @ Aq[q] @ x[sl]
    return g


def eigen_selfadjoint(G):
    L = eigen_lib()
    if L is None:
        return None
    n = G.shape[0]
    dp = ctypes.POINTER(ctypes.c_double)
    Gc = np.asfortranarray(G)
    ev = np.zeros(n)
    V = np.zeros((n, n), order="F")
    L.ref_selfadjoint_eig(n, Gc.ctypes.data_as(dp), ev.ctypes.data_as(dp), V.ctypes.data_as(dp))
    return ev, V


def eigen_bdcsvd(S):
    L = eigen_lib()
    if L is None:
        return None
    m, k = S.shape
    dp = ctypes.POINTER(ctypes.c_double)
    Sc = np.asfortranarray(S)
    sv = np.zeros(min(m, k))
    U = np.zeros((m, min(m, k)), order="F")
    L.ref_bdcsvd(m, k, Sc.ctypes.data_as(dp), sv.ctypes.data_as(dp), U.ctypes.data_as(dp))
    return sv, U
