"""POD offline stage on the GPU: snapshot Gram matrix, FP64 eigen-decomposition, basis,
Galerkin projection of matrix-free operator terms.

The reference builds its reduced basis from an Eigen BDCSVD of the raw snapshot matrix
and truncates by cumulative singular values (CoupledLinearNS_ROM.cpp:3916-3954), then
projects operator terms densely (RB^T A RB, :2788-2802, :3661-3663).  Here (method of
snapshots, mass-weighted):
    G = S^T M S                      (M = IProduct o BwdTrans, applied matrix-free)
    G = Psi Lambda Psi^T             (FP64 Jacobi on the GPU)
    V = S Psi_r Lambda_r^{-1/2}      (M-orthonormal modes, V^T M V = I)
    A_q^r = V^T (A_q V)              (A_q applied matrix-free, products on MFMA)
Storage convention: a set of k vectors of length N is a (k, N) row-major tensor, i.e. an
N x k column-major matrix for the C ABI.
"""
import ctypes

import numpy as np
import torch

from . import MFError
from . import rom as _rom

GRAM_FP64, GRAM_FP32X3, GRAM_BF16X3 = 0, 1, 2
_SET = False


def _lib():
    global _SET
    L = _rom.lib()
    if not _SET:
        vp, i, l_, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_size_t
        L.irom_gram_workspace_size.restype = sz
        L.irom_gram_workspace_size.argtypes = [i, i, l_]
        L.irom_gram.argtypes = [i, i, i, l_, vp, l_, vp, l_, vp, l_, i, vp, sz, vp]
        L.irom_gram_last_error.restype = ctypes.c_char_p
        L.irom_eig_sym.argtypes = [i, vp, vp, vp, d, i, ctypes.POINTER(i), vp]
        L.irom_eig_last_error.restype = ctypes.c_char_p
        _SET = True
    return L


def _s(stream=None):
    return ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


def gram(X, Y=None, mode=GRAM_FP64, out=None, accumulate=False, work=None, stream=None):
    """C = X Y^T for row-major X (m, K) and Y (n, K) (= X^T Y of the column-major N x k
    matrices).  X with Y=None computes the symmetric X X^T."""
    L = _lib()
    sym = Y is None
    Y = X if sym else Y
    m, K = X.shape[0], X.shape[1]
    n = Y.shape[0]
    assert Y.shape[1] == K and X.dtype == torch.float64 and Y.dtype == torch.float64
    assert X.stride(1) == 1 and Y.stride(1) == 1
    if out is None:
        out = torch.zeros((m, n), dtype=torch.float64, device=X.device)
    need = L.irom_gram_workspace_size(m, n, K)
    if work is None or work.numel() * 8 < need:
        work = torch.empty((need + 7) // 8, dtype=torch.float64, device=X.device)
    st = L.irom_gram(mode, m, n, K, ctypes.c_void_p(X.data_ptr()), X.stride(0),
                     ctypes.c_void_p(Y.data_ptr()), Y.stride(0), ctypes.c_void_p(out.data_ptr()),
                     out.stride(0), int(accumulate), ctypes.c_void_p(work.data_ptr()),
                     work.numel() * 8, _s(stream))
    if st:
        raise MFError(L.irom_gram_last_error().decode())
    return out


def eig_sym(G, tol=None, max_sweeps=60, stream=None):
    """Eigenpairs of a symmetric PSD matrix (descending).  Returns (evals numpy, V) with
    V[j] the j-th eigenvector (row j of the returned (n, n) tensor), and the sweep count."""
    L = _lib()
    n = G.shape[0]
    if tol is None:
        tol = max(1e-14, n * 2.220446049250313e-16)
    G = G.contiguous()
    ev = np.zeros(n)
    V = torch.empty((n, n), dtype=torch.float64, device=G.device)
    sw = ctypes.c_int()
    st = L.irom_eig_sym(n, ctypes.c_void_p(G.data_ptr()), ev.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_void_p(V.data_ptr()), tol, max_sweeps, ctypes.byref(sw), _s(stream))
    if st:
        raise MFError(L.irom_eig_last_error().decode())
    return ev, V, sw.value


def energy_truncation(evals, retained=1.0 - 1e-8):
    """Smallest r with sum_{i<r} lambda_i >= retained * sum lambda (lambda = sigma^2 of
    M^{1/2} S)."""
    lam = np.clip(np.asarray(evals), 0.0, None)
    c = np.cumsum(lam) / lam.sum()
    return int(np.searchsorted(c, retained) + 1)


def reference_rbsize(sigma, pod_tolerance):
    """The reference's rule on singular values sigma (CoupledLinearNS_ROM.cpp:3934-3946):
    RBsize = 1, raised to i+2 for every i with sum_{<=i} sigma / sum sigma < tolerance."""
    s = np.asarray(sigma)
    c = np.cumsum(s) / s.sum()
    rb = 1
    for i in range(len(s)):
        if c[i] < pod_tolerance:
            rb = i + 2
    return rb


def mass_apply(S, bwd, ipr, out=None, phys=None):
    """M s_j = IProduct(BwdTrans(s_j)) for every row of S (k, N)."""
    if out is None:
        out = torch.empty_like(S)
    if phys is None:
        phys = torch.empty(bwd.out_size, dtype=torch.float64, device=S.device)
    for j in range(S.shape[0]):
        bwd.apply(S[j], phys)
        ipr.apply(phys, out[j])
    return out


def pod_basis(S, evals, evecs, r):
    """V = Psi_r^T-weighted combination of the snapshots: V (r, N) = Lambda_r^{-1/2}
    Psi_r^T S with Psi_r the leading r eigenvectors (rows of evecs)."""
    C = evecs[:r] / torch.sqrt(torch.as_tensor(evals[:r], device=S.device))[:, None]
    return C @ S
