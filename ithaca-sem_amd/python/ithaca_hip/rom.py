"""Reduced-basis online stage on the GPU: batched affine assembly and solves.

Python binding of include/ithaca_rom.h.  The reference solves one parameter at a time
(CoupledLinearNS_ROM.cpp: gen_affine_mat_proj :4694-4704, colPivHouseholderQr().solve
:4053); ``Batch.solve`` solves thousands of instances of uneven size in one launch.
"""
import ctypes

import numpy as np
import torch

from . import LIB_PATH, MFError

LU_FP64, LU_FP32_IR, LU_FP32_IR_FALLBACK = 0, 1, 2
KINDS = ("normalised", "literal", "nonsymmetric")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        import os
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.irom_batch_create.argtypes = [i, vp, vp, ctypes.POINTER(vp)]
        L.irom_batch_destroy.argtypes = [vp]
        L.irom_batch_layout.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.irom_batch_solve.argtypes = [vp, i, i, vp, vp, vp, vp, d, i, vp, vp, vp]
        L.irom_batch_num_variants.restype = i
        L.irom_batch_variant_name.restype = ctypes.c_char_p
        L.irom_batch_variant_name.argtypes = [i]
        L.irom_batch_set_variant.argtypes = [vp, i, i]
        L.irom_batch_autotune.argtypes = [vp, i, vp, vp, vp, vp, i, ctypes.POINTER(i), vp]
        L.irom_batch_capture.argtypes = [vp, i, i, vp, vp, vp, vp, d, i, vp, vp]
        L.irom_batch_replay.argtypes = [vp, vp]
        L.irom_batch_param_grad.argtypes = [vp, vp, vp, vp, vp, vp]
        L.irom_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(st):
    if st != 0:
        raise MFError(lib().irom_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(s):
    return ctypes.c_void_p((s or torch.cuda.current_stream()).cuda_stream)


class Batch:
    """A batch of affine reduced systems A(mu_i) x_i = b_i, A(mu) = sum_q theta_q A_q."""

    def __init__(self, N, Q):
        self.N = np.ascontiguousarray(N, dtype=np.int32)
        self.Q = np.ascontiguousarray(Q, dtype=np.int32)
        self.n = len(self.N)
        h = ctypes.c_void_p()
        _check(lib().irom_batch_create(self.n, self.N.ctypes.data_as(ctypes.c_void_p),
                                       self.Q.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self._h = h
        self.a_off = np.zeros(self.n, dtype=np.int64)
        self.t_off = np.zeros(self.n, dtype=np.int64)
        self.v_off = np.zeros(self.n, dtype=np.int64)
        sizes = (ctypes.c_size_t * 3)()
        _check(lib().irom_batch_layout(self._h, self.a_off.ctypes.data_as(ctypes.c_void_p),
                                       self.t_off.ctypes.data_as(ctypes.c_void_p),
                                       self.v_off.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.byref(sizes, 0), ctypes.byref(sizes, 8),
                                       ctypes.byref(sizes, 16)))
        self.a_size, self.t_size, self.v_size = int(sizes[0]), int(sizes[1]), int(sizes[2])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _LIB is not None:
            _LIB.irom_batch_destroy(h)
            self._h = None

    def solve(self, A, theta, b, x=None, method=LU_FP64, transpose=False, tol=1e-14,
              max_iter=10, iters=None, status=None, stream=None):
        """x_i = A(mu_i)^{-1} b_i (or A^{-T} b_i) for every instance; cuda float64 tensors."""
        for t, n in ((A, self.a_size), (theta, self.t_size), (b, self.v_size)):
            assert t.is_cuda and t.dtype == torch.float64 and t.numel() == n and t.is_contiguous()
        if x is None:
            x = torch.empty_like(b)
        _check(lib().irom_batch_solve(self._h, method, int(transpose), _ptr(A), _ptr(theta),
                                      _ptr(b), _ptr(x), float(tol), int(max_iter), _ptr(iters),
                                      _ptr(status), _stream(stream)))
        return x

    def variants(self):
        return [lib().irom_batch_variant_name(v).decode()
                for v in range(lib().irom_batch_num_variants())]

    def set_variant(self, method, v):
        _check(lib().irom_batch_set_variant(self._h, method, v))

    def autotune(self, A, theta, b, x, method, reps=2, stream=None):
        c = ctypes.c_int()
        _check(lib().irom_batch_autotune(self._h, method, _ptr(A), _ptr(theta), _ptr(b), _ptr(x),
                                         reps, ctypes.byref(c), _stream(stream)))
        return c.value

    def capture(self, A, theta, b, x, method=LU_FP64, transpose=False, tol=1e-14, max_iter=10,
                iters=None, status=None):
        _check(lib().irom_batch_capture(self._h, method, int(transpose), _ptr(A), _ptr(theta),
                                        _ptr(b), _ptr(x), float(tol), int(max_iter), _ptr(iters),
                                        _ptr(status)))

    def replay(self, stream=None):
        _check(lib().irom_batch_replay(self._h, _stream(stream)))

    def param_grad(self, A, lam, x, grad=None, stream=None):
        """grad[t_off[i] + q] = -lam_i^T A_q x_i (dJ/dmu_q for J = l^T x, A^T lam = l)."""
        if grad is None:
            grad = torch.empty(self.t_size, dtype=torch.float64, device=A.device)
        _check(lib().irom_batch_param_grad(self._h, _ptr(A), _ptr(lam), _ptr(x), _ptr(grad),
                                           _stream(stream)))
        return grad


def _layout(rng, n_inst, nmin, nmax, qmin, qmax):
    Q = rng.integers(qmin, qmax + 1, size=n_inst)
    N = rng.integers(nmin, nmax + 1, size=n_inst)
    a_off = np.zeros(n_inst, dtype=np.int64)
    t_off = np.zeros(n_inst, dtype=np.int64)
    v_off = np.zeros(n_inst, dtype=np.int64)
    a_off[1:] = np.cumsum(Q * N * N)[:-1]
    t_off[1:] = np.cumsum(Q)[:-1]
    v_off[1:] = np.cumsum(N)[:-1]
    return Q, N, a_off, t_off, v_off


def affine_batch(n_inst=4096, seed=3, nmin=16, nmax=256, qmin=3, qmax=8, kind="normalised"):
    """Synthetic affine reduced systems (host numpy, float64).

    Per instance: Q ~ U{qmin..qmax} terms and size N ~ U{nmin..nmax}; theta_q = mu_q ~
    U[0.1, 1]; b ~ N(0, 1); A_0 built from Haar-orthogonal factors (QR of a Gaussian
    matrix, signs fixed by diag(R)) and singular values 10^{U[0,4]}.  ``kind``:

    * ``"normalised"``   A_0 = U diag U^T, A_q (q >= 1) = 0.1 sym(G) / sqrt(2N) with
      sym(G) = (G + G^T)/2: symmetric terms scaled to spectral norm ~0.1 (sqrt(2N) is the
      semicircle edge of sym(G)), so every A(mu) is symmetric positive definite with
      condition number ~1e4;
    * ``"literal"``      A_0 = U diag U^T, A_q = 0.1 sym(G) unnormalised, the workload as
      stated: about 38 % of the A(mu) are indefinite and condition numbers reach ~1e7;
    * ``"nonsymmetric"`` A_0 = U diag V^T with independent U, V, and A_q = 0.05 G / sqrt(N)
      (spectral norm ~0.1): non-symmetric systems, as the reference's coupled
      Navier-Stokes reduced matrices are.

    Returns the packed arrays of include/ithaca_rom.h in instance order."""
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {KINDS}")
    rng = np.random.default_rng(seed)
    Q, N, a_off, t_off, v_off = _layout(rng, n_inst, nmin, nmax, qmin, qmax)
    A = np.empty(int(np.sum(Q * N * N)))
    theta = rng.uniform(0.1, 1.0, size=int(Q.sum()))
    b = rng.standard_normal(int(N.sum()))
    for n in np.unique(N):
        idx = np.nonzero(N == n)[0]
        G = rng.standard_normal((len(idx), n, n))
        U, R = np.linalg.qr(G)
        U *= np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
        lam = 10.0 ** rng.uniform(0.0, 4.0, size=(len(idx), n))
        if kind == "nonsymmetric":
            V, R = np.linalg.qr(rng.standard_normal((len(idx), n, n)))
            V *= np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
        else:
            V = U
        A0 = (U * lam[:, None, :]) @ np.swapaxes(V, 1, 2)
        npert = int(np.sum(Q[idx] - 1))
        P = rng.standard_normal((npert, n, n))
        if kind == "normalised":
            P = (0.05 / np.sqrt(2.0 * n)) * (P + np.swapaxes(P, 1, 2))
        elif kind == "literal":
            P = 0.05 * (P + np.swapaxes(P, 1, 2))
        else:
            P *= 0.05 / np.sqrt(n)
        c = 0
        for r, i in enumerate(idx):
            base = a_off[i]
            A[base:base + n * n] = A0[r].ravel()
            k = Q[i] - 1
            A[base + n * n: base + Q[i] * n * n] = P[c:c + k].ravel()
            c += k
    return dict(N=N.astype(np.int32), Q=Q.astype(np.int32), A=A, theta=theta, b=b,
                a_off=a_off, t_off=t_off, v_off=v_off)


def affine_batch_device(n_inst=4096, seed=3, nmin=16, nmax=256, qmin=3, qmax=8,
                        device="cuda", kind="normalised"):
    """The distributions of ``affine_batch`` generated on the GPU (for large batches).

    N, Q, theta and b are drawn exactly as in ``affine_batch`` (numpy PCG64, same seed);
    the matrices come from torch's Philox generator seeded with ``seed`` on the device, so
    their values differ from the host generator's.  Returns the same dict with ``A`` a cuda
    float64 tensor and ``theta``, ``b`` as host arrays."""
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {KINDS}")
    rng = np.random.default_rng(seed)
    Q, N, a_off, t_off, v_off = _layout(rng, n_inst, nmin, nmax, qmin, qmax)
    theta = rng.uniform(0.1, 1.0, size=int(Q.sum()))
    b = rng.standard_normal(int(N.sum()))
    gen = torch.Generator(device=device).manual_seed(seed)
    A = torch.empty(int(np.sum(Q * N * N)), dtype=torch.float64, device=device)

    def haar(m, n):
        G = torch.randn((m, n, n), generator=gen, dtype=torch.float64, device=device)
        U, R = torch.linalg.qr(G)
        return U * torch.sign(torch.diagonal(R, dim1=1, dim2=2))[:, None, :]

    for n in np.unique(N):
        n = int(n)
        idx = np.nonzero(N == n)[0]
        U = haar(len(idx), n)
        lam = 10.0 ** (4.0 * torch.rand((len(idx), n), generator=gen, dtype=torch.float64,
                                        device=device))
        V = haar(len(idx), n) if kind == "nonsymmetric" else U
        A0 = (U * lam[:, None, :]) @ V.transpose(1, 2).contiguous()
        npert = int(np.sum(Q[idx] - 1))
        P = torch.randn((npert, n, n), generator=gen, dtype=torch.float64, device=device)
        if kind == "normalised":
            P = (0.05 / np.sqrt(2.0 * n)) * (P + P.transpose(1, 2))
        elif kind == "literal":
            P = 0.05 * (P + P.transpose(1, 2))
        else:
            P = (0.05 / np.sqrt(n)) * P
        c = 0
        for r, i in enumerate(idx):
            base = int(a_off[i])
            A[base:base + n * n] = A0[r].reshape(-1)
            k = int(Q[i]) - 1
            A[base + n * n: base + (k + 1) * n * n] = P[c:c + k].reshape(-1)
            c += k
    return dict(N=N.astype(np.int32), Q=Q.astype(np.int32), A=A, theta=theta, b=b,
                a_off=a_off, t_off=t_off, v_off=v_off)
