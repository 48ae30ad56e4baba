"""CPU oracle for the ITHACA-SEM matrix-free operator path (TEST INFRASTRUCTURE).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import it.

* ``liboracle_mf.so`` (built from ``mf_oracle.c`` by ``oracle/Makefile``) restates the
  reference CPU algorithms: Polylib GLL tables, the Modified_A basis, trilinear
  geometric factors and the MatrixFree BwdTrans / IProduct / PhysDeriv / Helmholtz
  kernels for quads and hexes in the reference's summation order.
* ``mesh`` builds the synthetic meshes and inputs of the measured workloads.
* ``rom`` holds the FP64 numpy/LAPACK restatement of the reduced-basis kernels.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle_mf.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle_mf.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().or_set_threads(ctypes.c_int(int(n)))


def basis(nm, nq):
    """Modified_A basis on nq GLL points: dict(z, w, D, bdata, dbdata)."""
    z = np.zeros(nq); w = np.zeros(nq); D = np.zeros(nq * nq)
    b = np.zeros(nm * nq); db = np.zeros(nm * nq)
    lib().or_basis_modified_a(nm, nq, _p(z), _p(w), _p(D), _p(b), _p(db))
    return dict(z=z, w=w, D=D, bdata=b, dbdata=db, nm=nm, nq=nq)


def geom_hex(verts, nq, z, deformed=True):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    Q = nq ** 3
    n = E * Q if deformed else E
    jac = np.zeros(n); df = np.zeros(9 * n)
    lib().or_geom_hex(_p(verts), E, nq, _p(z), int(deformed), _p(jac), _p(df))
    return jac, df


def geom_quad(verts, nq, z, deformed=True):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    Q = nq ** 2
    n = E * Q if deformed else E
    jac = np.zeros(n); df = np.zeros(4 * n)
    lib().or_geom_quad(_p(verts), E, nq, _p(z), int(deformed), _p(jac), _p(df))
    return jac, df


def coords_hex(verts, nq, z):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xyz = np.zeros(3 * E * nq ** 3)
    lib().or_coords_hex(_p(verts), E, nq, _p(z), _p(xyz))
    return xyz.reshape(3, -1)


def coords_quad(verts, nq, z):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xy = np.zeros(2 * E * nq ** 2)
    lib().or_coords_quad(_p(verts), E, nq, _p(z), _p(xy))
    return xy.reshape(2, -1)


def bwdtrans(shape, b, E, u):
    nm, nq = b["nm"], b["nq"]
    d = 3 if shape == "hex" else 2
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * nq ** d)
    fn = lib().or_bwdtrans_hex if d == 3 else lib().or_bwdtrans_quad
    fn(nm, nq, _p(b["bdata"]), E, _p(u), _p(out))
    return out


def iproduct(shape, b, E, deformed, jac, f):
    nm, nq = b["nm"], b["nq"]
    d = 3 if shape == "hex" else 2
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * nm ** d)
    fn = lib().or_iproduct_hex if d == 3 else lib().or_iproduct_quad
    fn(nm, nq, _p(b["bdata"]), _p(b["w"]), E, int(deformed), _p(jac), _p(f), _p(out))
    return out


def physderiv(shape, b, E, deformed, df, f):
    nq = b["nq"]
    d = 3 if shape == "hex" else 2
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(E * nq ** d) for _ in range(d)]
    if d == 3:
        lib().or_physderiv_hex(nq, _p(b["D"]), E, int(deformed), _p(df), _p(f),
                               *[_p(o) for o in outs])
    else:
        lib().or_physderiv_quad(nq, _p(b["D"]), E, int(deformed), _p(df), _p(f),
                                *[_p(o) for o in outs])
    return outs


def iprod_deriv(shape, b, E, deformed, jac, df, ins):
    """IProductWRTDerivBase: ins = dim arrays of E*Q point values -> E*M coefficients."""
    nm, nq = b["nm"], b["nq"]
    d = 3 if shape == "hex" else 2
    ins = [np.ascontiguousarray(x, dtype=np.float64) for x in ins]
    out = np.zeros(E * nm ** d)
    if d == 3:
        lib().or_iprod_deriv_hex(nm, nq, _p(b["bdata"]), _p(b["dbdata"]), _p(b["w"]), E,
                                 int(deformed), _p(jac), _p(df), *[_p(x) for x in ins],
                                 _p(out))
    else:
        lib().or_iprod_deriv_quad(nm, nq, _p(b["bdata"]), _p(b["dbdata"]), _p(b["w"]), E,
                                  int(deformed), _p(jac), _p(df), *[_p(x) for x in ins],
                                  _p(out))
    return out


def helmholtz(shape, b, E, deformed, jac, df, lam, u, simd="scalar"):
    """simd: "scalar" (width 1, the reference's default build) or "avx2" (hexes: four
    elements interleaved per vector, the reference's AVX2 SIMD ABI)."""
    nm, nq = b["nm"], b["nq"]
    d = 3 if shape == "hex" else 2
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * nm ** d)
    fn = lib().or_helmholtz_hex if d == 3 else lib().or_helmholtz_quad
    if simd == "avx2":
        if d != 3:
            raise ValueError("the AVX2 variant covers hexahedra")
        fn = lib().or_helmholtz_hex_avx2
    fn(nm, nq, _p(b["bdata"]), _p(b["dbdata"]), _p(b["D"]), _p(b["w"]), E, int(deformed),
       _p(jac), _p(df), ctypes.c_double(lam), _p(u), _p(out))
    return out
