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

from . import FP64, MFError, Operator
from . import rom as _rom

GRAM_FP64, GRAM_FP32X3, GRAM_BF16X3 = 0, 1, 2
GRAM_SYMMETRIC_RESULT = 16
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
        L.irom_galerkin_project.argtypes = [vp, i, vp, l_, vp, l_, vp, l_, i, vp, sz, vp]
        L.irom_project_last_error.restype = ctypes.c_char_p
        _SET = True
    return L


def _s(stream=None):
    return ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


def gram(X, Y=None, mode=GRAM_FP64, out=None, accumulate=False, work=None, stream=None,
         symmetric_result=False):
    """C = X Y^T for row-major X (m, K) and Y (n, K) (= X^T Y of the column-major N x k
    matrices).  X with Y=None computes the symmetric X X^T; symmetric_result=True states
    that X Y^T is symmetric (e.g. S (M S)^T with M symmetric): only the upper blocks are
    computed and mirrored."""
    if symmetric_result:
        mode |= GRAM_SYMMETRIC_RESULT
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


def galerkin_project(op, V, mode=GRAM_FP64, W=None, out=None, work=None, stream=None,
                     symmetric_result=True):
    """C = V A V^T (r x r) for the basis V (r, N) and the matrix-free operator term A
    (Helmholtz or Mass, FP64): A applied to all r rows in one launch (into W, returned
    too), then the product on MFMA in `mode`.  The reference's per-term RB^T A RB
    (CoupledLinearNS_ROM.cpp:2788-2802)."""
    L = _lib()
    r, N = V.shape
    assert V.dtype == torch.float64 and V.stride(1) == 1
    if W is None:
        W = torch.empty((r, N), dtype=torch.float64, device=V.device)
    if out is None:
        out = torch.empty((r, r), dtype=torch.float64, device=V.device)
    if symmetric_result:
        mode |= GRAM_SYMMETRIC_RESULT
    need = L.irom_gram_workspace_size(r, r, N)
    if work is None or work.numel() * 8 < need:
        work = torch.empty((need + 7) // 8, dtype=torch.float64, device=V.device)
    st = L.irom_galerkin_project(op._h, r, ctypes.c_void_p(V.data_ptr()), V.stride(0),
                                 ctypes.c_void_p(W.data_ptr()), W.stride(0),
                                 ctypes.c_void_p(out.data_ptr()), out.stride(0), mode,
                                 ctypes.c_void_p(work.data_ptr()), work.numel() * 8, _s(stream))
    if st:
        raise MFError(L.irom_project_last_error().decode())
    return out, W


def project_affine(terms, V, mode=GRAM_FP64):
    """Reduced matrices of the affine terms {name: operator} on the basis V (r, N), one
    Galerkin projection per term; the reduced operator of a parameter is
    sum_q theta_q C_q (gen_affine_mat_proj, CoupledLinearNS_ROM.cpp:4694-4704)."""
    W = torch.empty_like(V)
    return {name: galerkin_project(op, V, mode, W=W)[0].clone() for name, op in terms.items()}


def helmholtz_affine_terms(nm, nq, nElmt, basis, jac, df, deformed=True):
    """The two affine terms of the hex Helmholtz operator A(lambda) = K + lambda M as
    matrix-free operators: the stiffness K (Helmholtz with lambda = 0) and the mass M."""
    geo = "Deformed" if deformed else "Regular"
    K = Operator(f"Helmholtz_Hex_{geo}", nm, nq, nElmt, FP64, basis)
    K.set_jac(jac)
    K.set_df(df)
    K.set_lambda(0.0)
    Mo = Operator(f"Mass_Hex_{geo}", nm, nq, nElmt, FP64, basis)
    Mo.set_jac(jac)
    return {"stiffness": K, "mass": Mo}


def mass_apply(S, bwd, ipr, out=None, phys=None):
    """M s_j = IProduct(BwdTrans(s_j)) for every row of S (k, N), one pair of operator
    applies per row (the reference's MassMatrixOp_MatFree, StdExpansion.cpp:647-661)."""
    if out is None:
        out = torch.empty_like(S)
    if phys is None:
        phys = torch.empty(bwd.out_size, dtype=torch.float64, device=S.device)
    for j in range(S.shape[0]):
        bwd.apply(S[j], phys)
        ipr.apply(phys, out[j])
    return out


def mass_apply_fused(S, mass, out=None, block=64):
    """M s_j for every row of S (k, N) with the fused "Mass_Hex_*" operator: one launch
    per block of rows, the geometry read once per launch."""
    if out is None:
        out = torch.empty_like(S)
    for j0 in range(0, S.shape[0], block):
        j1 = min(S.shape[0], j0 + block)
        mass.apply_many(S[j0:j1], out[j0:j1])
    return out


def pod_basis(S, evals, evecs, r):
    """V = Psi_r^T-weighted combination of the snapshots: V (r, N) = Lambda_r^{-1/2}
    Psi_r^T S with Psi_r the leading r eigenvectors (rows of evecs)."""
    C = evecs[:r] / torch.sqrt(torch.as_tensor(evals[:r], device=S.device))[:, None]
    return C @ S


class ElementalProjector:
    """Elemental L2 projection onto the hex space, u_e = M_e^{-1} IProduct_e(f), for
    straight-sided deformed hexes: the reference's per-element FwdTrans
    (LocalRegions/HexExp.cpp:325-349, IProductWRTBase then the element's eInvMass).

    M_e = B^T diag(J w) B is applied matrix-free (the fused Mass operator).  The
    system is solved by conjugate gradients on the block-diagonal mass matrix,
    preconditioned by the element's constant-Jacobian inverse
    (M1^{-1} x M1^{-1} x M1^{-1}) / Jbar_e, which is exact when J is constant on the
    element (M1 = 1D mass matrix, Jbar_e = weighted mean of J); its Kronecker product is
    applied by the sum-factorisation BwdTrans kernel with M1^{-1} as the table."""

    def __init__(self, nElmt, nm, nq, basis, jac):
        self.E, self.nm, self.nq = nElmt, nm, nq
        self.M, self.Q = nm ** 3, nq ** 3
        self.bwd = Operator("BwdTrans_Hex_Regular", nm, nq, nElmt, FP64, basis)
        self.ipr = Operator("IProduct_Hex_Deformed", nm, nq, nElmt, FP64, basis)
        self.ipr.set_jac(jac)
        self.mass_op = Operator("Mass_Hex_Deformed", nm, nq, nElmt, FP64, basis)
        self.mass_op.set_jac(jac)
        B = np.asarray(basis["bdata"]).reshape(nm, nq)
        w = np.asarray(basis["w"])
        Minv = np.linalg.inv((B * w) @ B.T)
        kb = dict(bdata=np.ascontiguousarray(Minv.ravel()), dbdata=np.zeros(nm * nm),
                  D=np.zeros(nm * nm), w=np.ones(nm), nm=nm, nq=nm)
        self.pre = Operator("BwdTrans_Hex_Regular", nm, nm, nElmt, FP64, kb)
        w3 = torch.as_tensor(np.einsum("k,j,i->kji", w, w, w).ravel(), device=jac.device)
        jbar = (jac.view(nElmt, self.Q) * w3).sum(1) / w3.sum()
        self.inv_jbar = (1.0 / jbar)[:, None].contiguous()
        self.phys = torch.empty(nElmt * self.Q, dtype=torch.float64, device=jac.device)

    def mass(self, x, out=None):
        return self.mass_op.apply(x, out)

    def precond(self, r, out=None):
        out = self.pre.apply(r, out)
        out.view(self.E, self.M).mul_(self.inv_jbar)
        return out

    def project(self, f, tol=1e-13, max_iter=200, check_every=4):
        """Coefficients of the elemental L2 projection of the point values f (E*Q);
        returns (u, iterations, relative residual ||b - M u|| / ||b||)."""
        b = self.ipr.apply(f)
        bn = torch.linalg.vector_norm(b)
        x = self.precond(b)
        r = b - self.mass(x)
        z = self.precond(r)
        p = z.clone()
        rz = torch.dot(r, z)
        Ap = torch.empty_like(b)
        it = 0
        for it in range(1, max_iter + 1):
            self.mass(p, Ap)
            alpha = rz / torch.dot(p, Ap)
            x.add_(alpha * p)
            r.sub_(alpha * Ap)
            if it % check_every == 0 and float(torch.linalg.vector_norm(r) / bn) <= tol:
                break
            self.precond(r, z)
            rz_new = torch.dot(r, z)
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
        res = float(torch.linalg.vector_norm(b - self.mass(x)) / bn)
        return x, it, res
