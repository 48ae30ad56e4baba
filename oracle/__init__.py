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


# ---- triangles (mf_oracle_tri.c) ----

def basis_tri(nm, nq0, nq1=None):
    """Modified_A on nq0 GLL points (eta0) and Modified_B on nq1 Gauss-Radau (alpha = 1)
    points (eta1, default nq0 - 1 as MatrixFree requires); w1 is the Helper's 0.5-scaled
    weight, w1_raw the Gauss-Radau one."""
    nq1 = nq0 - 1 if nq1 is None else nq1
    a = basis(nm, nq0)
    M = nm * (nm + 1) // 2
    z1 = np.zeros(nq1); w1 = np.zeros(nq1); D1 = np.zeros(nq1 * nq1)
    b1 = np.zeros(M * nq1); db1 = np.zeros(M * nq1)
    lib().or_basis_modified_b(nm, nq1, _p(z1), _p(w1), _p(D1), _p(b1), _p(db1))
    return dict(nm=nm, nq0=nq0, nq1=nq1, M=M, Q=nq0 * nq1,
                z0=a["z"], w0=a["w"], D0=a["D"], bdata0=a["bdata"], dbdata0=a["dbdata"],
                z1=z1, w1_raw=w1, w1=0.5 * w1, D1=D1, bdata1=b1, dbdata1=db1)


def geom_tri(verts, b, deformed=False):
    """Straight-sided triangles (E x 3 x 2 vertices): jac and GetDerivFactors layouts."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    n = E * b["Q"] if deformed else E
    jac = np.zeros(n); df = np.zeros(4 * n)
    lib().or_geom_tri(_p(verts), E, b["nq0"], b["nq1"], int(deformed), _p(jac), _p(df))
    return jac, df


def coords_tri(verts, b):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xy = np.zeros(2 * E * b["Q"])
    lib().or_coords_tri(_p(verts), E, b["nq0"], b["nq1"], _p(b["z0"]), _p(b["z1"]), _p(xy))
    return xy.reshape(2, -1)


def bwdtrans_tri(b, E, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["Q"])
    lib().or_bwdtrans_tri(b["nm"], b["nq0"], b["nq1"], _p(b["bdata0"]), _p(b["bdata1"]), E,
                          _p(u), _p(out))
    return out


def iproduct_tri(b, E, deformed, jac, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_iproduct_tri(b["nm"], b["nq0"], b["nq1"], _p(b["bdata0"]), _p(b["bdata1"]),
                          _p(b["w0"]), _p(b["w1"]), E, int(deformed), _p(jac), _p(f), _p(out))
    return out


def physderiv_tri(b, E, deformed, df, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(E * b["Q"]) for _ in range(2)]
    lib().or_physderiv_tri(b["nq0"], b["nq1"], _p(b["D0"]), _p(b["D1"]), _p(b["z0"]),
                           _p(b["z1"]), E, int(deformed), _p(df), _p(f), *[_p(o) for o in outs])
    return outs


def helmholtz_tri(b, E, deformed, jac, df, lam, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_helmholtz_tri(b["nm"], b["nq0"], b["nq1"], _p(b["bdata0"]), _p(b["dbdata0"]),
                           _p(b["bdata1"]), _p(b["dbdata1"]), _p(b["D0"]), _p(b["D1"]),
                           _p(b["z0"]), _p(b["z1"]), _p(b["w0"]), _p(b["w1"]), E,
                           int(deformed), _p(jac), _p(df), ctypes.c_double(lam), _p(u),
                           _p(out))
    return out


# ---- prisms (mf_oracle_prism.c) ----

def basis_prism(nm, nq, nq2=None):
    """Modified_A on nq GLL points along eta0 and xi1, Modified_B on nq2 Gauss-Radau
    (alpha = 1) points along eta2 (default nq - 1, as MatrixFree requires); w2 is the
    Helper's 0.5-scaled weight."""
    nq2 = nq - 1 if nq2 is None else nq2
    a = basis(nm, nq)
    t = basis_tri(nm, nq, nq2)
    return dict(nm=nm, nq0=nq, nq1=nq, nq2=nq2, M=nm * nm * (nm + 1) // 2, Q=nq * nq * nq2,
                z0=a["z"], w0=a["w"], D0=a["D"], bdata0=a["bdata"], dbdata0=a["dbdata"],
                z1=a["z"], w1=a["w"], D1=a["D"], bdata1=a["bdata"], dbdata1=a["dbdata"],
                z2=t["z1"], w2_raw=t["w1_raw"], w2=t["w1"], D2=t["D1"], bdata2=t["bdata1"],
                dbdata2=t["dbdata1"])


def geom_prism(verts, b, deformed=True):
    """Straight-sided prisms (E x 6 x 3 vertices in StdPrismExp order)."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    n = E * b["Q"] if deformed else E
    jac = np.zeros(n); df = np.zeros(9 * n)
    lib().or_geom_prism(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], _p(b["z0"]), _p(b["z1"]),
                        _p(b["z2"]), int(deformed), _p(jac), _p(df))
    return jac, df


def coords_prism(verts, b):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xyz = np.zeros(3 * E * b["Q"])
    lib().or_coords_prism(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], _p(b["z0"]), _p(b["z1"]),
                          _p(b["z2"]), _p(xyz))
    return xyz.reshape(3, -1)


def bwdtrans_prism(b, E, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["Q"])
    lib().or_bwdtrans_prism(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                            _p(b["bdata1"]), _p(b["bdata2"]), E, _p(u), _p(out))
    return out


def iproduct_prism(b, E, deformed, jac, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_iproduct_prism(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                            _p(b["bdata1"]), _p(b["bdata2"]), _p(b["w0"]), _p(b["w1"]),
                            _p(b["w2"]), E, int(deformed), _p(jac), _p(f), _p(out))
    return out


def physderiv_prism(b, E, deformed, df, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(E * b["Q"]) for _ in range(3)]
    lib().or_physderiv_prism(b["nq0"], b["nq1"], b["nq2"], _p(b["D0"]), _p(b["D1"]), _p(b["D2"]),
                             _p(b["z0"]), _p(b["z2"]), E, int(deformed), _p(df), _p(f),
                             *[_p(o) for o in outs])
    return outs


def helmholtz_prism(b, E, deformed, jac, df, lam, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_helmholtz_prism(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                             _p(b["dbdata0"]), _p(b["bdata1"]), _p(b["dbdata1"]),
                             _p(b["bdata2"]), _p(b["dbdata2"]), _p(b["D0"]), _p(b["D1"]),
                             _p(b["D2"]), _p(b["z0"]), _p(b["z2"]), _p(b["w0"]), _p(b["w1"]),
                             _p(b["w2"]), E, int(deformed), _p(jac), _p(df),
                             ctypes.c_double(lam), _p(u), _p(out))
    return out


# ---- tetrahedra (mf_oracle_tet.c) ----

def basis_tet(nm, nq0, nq1=None, nq2=None):
    """Modified_A on nq0 GLL points (eta0), Modified_B on nq1 Gauss-Radau alpha = 1 points
    (eta1) and Modified_C on nq2 Gauss-Radau alpha = 2 points (eta2), the last two by
    default nq0 - 1 as MatrixFree requires; w1 / w2 are the Helper's 0.5- / 0.25-scaled
    weights, w1_raw / w2_raw the Gauss-Radau ones."""
    nq1 = nq0 - 1 if nq1 is None else nq1
    nq2 = nq1 if nq2 is None else nq2
    t = basis_tri(nm, nq0, nq1)
    M = nm * (nm + 1) * (nm + 2) // 6
    z2 = np.zeros(nq2); w2 = np.zeros(nq2); D2 = np.zeros(nq2 * nq2)
    b2 = np.zeros(M * nq2); db2 = np.zeros(M * nq2)
    lib().or_basis_modified_c(nm, nq2, _p(z2), _p(w2), _p(D2), _p(b2), _p(db2))
    return dict(nm=nm, nq0=nq0, nq1=nq1, nq2=nq2, M=M, Q=nq0 * nq1 * nq2,
                z0=t["z0"], w0=t["w0"], D0=t["D0"], bdata0=t["bdata0"], dbdata0=t["dbdata0"],
                z1=t["z1"], w1_raw=t["w1_raw"], w1=t["w1"], D1=t["D1"], bdata1=t["bdata1"],
                dbdata1=t["dbdata1"], z2=z2, w2_raw=w2, w2=0.25 * w2, D2=D2, bdata2=b2,
                dbdata2=db2)


def geom_tet(verts, b, deformed=False):
    """Straight-sided tetrahedra (E x 4 x 3 vertices in StdTetExp order)."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    n = E * b["Q"] if deformed else E
    jac = np.zeros(n); df = np.zeros(9 * n)
    lib().or_geom_tet(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], int(deformed), _p(jac),
                      _p(df))
    return jac, df


def coords_tet(verts, b):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xyz = np.zeros(3 * E * b["Q"])
    lib().or_coords_tet(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], _p(b["z0"]), _p(b["z1"]),
                        _p(b["z2"]), _p(xyz))
    return xyz.reshape(3, -1)


def bwdtrans_tet(b, E, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["Q"])
    lib().or_bwdtrans_tet(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                          _p(b["bdata1"]), _p(b["bdata2"]), E, _p(u), _p(out))
    return out


def iproduct_tet(b, E, deformed, jac, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_iproduct_tet(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                          _p(b["bdata1"]), _p(b["bdata2"]), _p(b["w0"]), _p(b["w1"]),
                          _p(b["w2"]), E, int(deformed), _p(jac), _p(f), _p(out))
    return out


def physderiv_tet(b, E, deformed, df, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(E * b["Q"]) for _ in range(3)]
    lib().or_physderiv_tet(b["nq0"], b["nq1"], b["nq2"], _p(b["D0"]), _p(b["D1"]), _p(b["D2"]),
                           _p(b["z0"]), _p(b["z1"]), _p(b["z2"]), E, int(deformed), _p(df),
                           _p(f), *[_p(o) for o in outs])
    return outs


def helmholtz_tet(b, E, deformed, jac, df, lam, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_helmholtz_tet(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                           _p(b["dbdata0"]), _p(b["bdata1"]), _p(b["dbdata1"]),
                           _p(b["bdata2"]), _p(b["dbdata2"]), _p(b["D0"]), _p(b["D1"]),
                           _p(b["D2"]), _p(b["z0"]), _p(b["z1"]), _p(b["z2"]), _p(b["w0"]),
                           _p(b["w1"]), _p(b["w2"]), E, int(deformed), _p(jac), _p(df),
                           ctypes.c_double(lam), _p(u), _p(out))
    return out


# ---- pyramids (mf_oracle_pyr.c) ----

def basis_pyr(nm, nq, nq2=None):
    """Modified_A on nq GLL points along eta0 and eta1, ModifiedPyr_C on nq2 Gauss-Radau
    (alpha = 2) points along eta2 (default nq - 1, as MatrixFree requires); w2 is the
    Helper's 0.25-scaled weight, w2_raw the Gauss-Radau one."""
    nq2 = nq - 1 if nq2 is None else nq2
    a = basis(nm, nq)
    M = nm * (nm + 1) * (2 * nm + 1) // 6
    z2 = np.zeros(nq2); w2 = np.zeros(nq2); D2 = np.zeros(nq2 * nq2)
    b2 = np.zeros(M * nq2); db2 = np.zeros(M * nq2)
    lib().or_basis_modified_pyr_c(nm, nq2, _p(z2), _p(w2), _p(D2), _p(b2), _p(db2))
    return dict(nm=nm, nq0=nq, nq1=nq, nq2=nq2, M=M, Q=nq * nq * nq2,
                z0=a["z"], w0=a["w"], D0=a["D"], bdata0=a["bdata"], dbdata0=a["dbdata"],
                z1=a["z"], w1=a["w"], D1=a["D"], bdata1=a["bdata"], dbdata1=a["dbdata"],
                z2=z2, w2_raw=w2, w2=0.25 * w2, D2=D2, bdata2=b2, dbdata2=db2)


def geom_pyr(verts, b, deformed=True):
    """Straight-sided pyramids (E x 5 x 3 vertices in StdPyrExp order)."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    n = E * b["Q"] if deformed else E
    jac = np.zeros(n); df = np.zeros(9 * n)
    lib().or_geom_pyr(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], _p(b["z0"]), _p(b["z1"]),
                      _p(b["z2"]), int(deformed), _p(jac), _p(df))
    return jac, df


def coords_pyr(verts, b):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E = verts.shape[0]
    xyz = np.zeros(3 * E * b["Q"])
    lib().or_coords_pyr(_p(verts), E, b["nq0"], b["nq1"], b["nq2"], _p(b["z0"]), _p(b["z1"]),
                        _p(b["z2"]), _p(xyz))
    return xyz.reshape(3, -1)


def bwdtrans_pyr(b, E, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["Q"])
    lib().or_bwdtrans_pyr(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                          _p(b["bdata1"]), _p(b["bdata2"]), E, _p(u), _p(out))
    return out


def iproduct_pyr(b, E, deformed, jac, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_iproduct_pyr(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                          _p(b["bdata1"]), _p(b["bdata2"]), _p(b["w0"]), _p(b["w1"]),
                          _p(b["w2"]), E, int(deformed), _p(jac), _p(f), _p(out))
    return out


def physderiv_pyr(b, E, deformed, df, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(E * b["Q"]) for _ in range(3)]
    lib().or_physderiv_pyr(b["nq0"], b["nq1"], b["nq2"], _p(b["D0"]), _p(b["D1"]), _p(b["D2"]),
                           _p(b["z0"]), _p(b["z1"]), _p(b["z2"]), E, int(deformed), _p(df),
                           _p(f), *[_p(o) for o in outs])
    return outs


def helmholtz_pyr(b, E, deformed, jac, df, lam, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["M"])
    lib().or_helmholtz_pyr(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]),
                           _p(b["dbdata0"]), _p(b["bdata1"]), _p(b["dbdata1"]),
                           _p(b["bdata2"]), _p(b["dbdata2"]), _p(b["D0"]), _p(b["D1"]),
                           _p(b["D2"]), _p(b["z0"]), _p(b["z1"]), _p(b["z2"]), _p(b["w0"]),
                           _p(b["w1"]), _p(b["w2"]), E, int(deformed), _p(jac), _p(df),
                           ctypes.c_double(lam), _p(u), _p(out))
    return out


# ---- IProductWRTDerivBase of the collapsed shapes ----

def iprodderiv_tri(b, E, deformed, jac, df, f0, f1):
    f0, f1 = (np.ascontiguousarray(x, dtype=np.float64) for x in (f0, f1))
    out = np.zeros(E * b["M"])
    lib().or_iprodderiv_tri(b["nm"], b["nq0"], b["nq1"], _p(b["bdata0"]), _p(b["dbdata0"]),
                            _p(b["bdata1"]), _p(b["dbdata1"]), _p(b["z0"]), _p(b["z1"]),
                            _p(b["w0"]), _p(b["w1"]), E, int(deformed), _p(jac), _p(df), _p(f0),
                            _p(f1), _p(out))
    return out


def _iprodderiv3(fn, b, E, deformed, jac, df, f, with_z1):
    f = [np.ascontiguousarray(x, dtype=np.float64) for x in f]
    out = np.zeros(E * b["M"])
    zs = [_p(b["z0"])] + ([_p(b["z1"])] if with_z1 else []) + [_p(b["z2"])]
    fn(b["nm"], b["nq0"], b["nq1"], b["nq2"], _p(b["bdata0"]), _p(b["dbdata0"]),
       _p(b["bdata1"]), _p(b["dbdata1"]), _p(b["bdata2"]), _p(b["dbdata2"]), *zs,
       _p(b["w0"]), _p(b["w1"]), _p(b["w2"]), E, int(deformed), _p(jac), _p(df),
       *[_p(x) for x in f], _p(out))
    return out


def iprodderiv_prism(b, E, deformed, jac, df, f):
    return _iprodderiv3(lib().or_iprodderiv_prism, b, E, deformed, jac, df, f, False)


def iprodderiv_tet(b, E, deformed, jac, df, f):
    return _iprodderiv3(lib().or_iprodderiv_tet, b, E, deformed, jac, df, f, True)


def iprodderiv_pyr(b, E, deformed, jac, df, f):
    return _iprodderiv3(lib().or_iprodderiv_pyr, b, E, deformed, jac, df, f, True)


# ------------------------------------------------------------------ segments

def geom_seg(verts, nq, deformed=False):
    """verts: (E, 2, ncoord) end points -> (jac, df) in the GetJac / GetDerivFactors layouts."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    E, ncoord = verts.shape[0], verts.shape[2]
    n = E * nq if deformed else E
    jac = np.zeros(n); df = np.zeros(ncoord * n)
    lib().or_geom_seg(_p(verts), E, ncoord, nq, int(deformed), _p(jac), _p(df))
    return jac, df


def bwdtrans_seg(b, E, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(E * b["nq"])
    lib().or_bwdtrans_seg(b["nm"], b["nq"], _p(b["bdata"]), E, _p(u), _p(out))
    return out


def iproduct_seg(b, E, deformed, jac, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["nm"])
    lib().or_iproduct_seg(b["nm"], b["nq"], _p(b["bdata"]), _p(b["w"]), E, int(deformed),
                          _p(jac), _p(f), _p(out))
    return out


def physderiv_seg(b, E, deformed, ncoord, df, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(ncoord * E * b["nq"])
    lib().or_physderiv_seg(b["nq"], _p(b["D"]), E, int(deformed), ncoord, _p(df), _p(f), _p(out))
    return out


def iprodderiv_seg(b, E, deformed, ncoord, jac, df, f):
    """f: ncoord arrays of E*nq point values back to back."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(E * b["nm"])
    lib().or_iprodderiv_seg(b["nm"], b["nq"], _p(b["dbdata"]), _p(b["w"]), E, int(deformed),
                            ncoord, _p(jac), _p(df), _p(f), _p(out))
    return out
