"""Pins of the oracle's IProductWRTDerivBase for the collapsed shapes (triangles, prisms,
tetrahedra, pyramids; MatrixFreeOps/IProductWRTDerivBase.h): against the dense
sum_c G_c^T diag(W J) f_c with the independent dense gradients of tests/dense_ref.py, and
the transpose identity <IPDB(f), u> = sum_c <f_c, W J d(B u)/dx_c> with the oracle's own
BwdTrans and PhysDeriv."""
import numpy as np
import pytest

import dense_ref as R
import oracle as O

TRI = np.array([[0.0, 0.0], [1.0, 0.2], [0.3, 1.0]])
PRISM = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
                  [0.5, 0.0, 1.0], [0.5, 1.0, 1.0]])
TET = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
PYR = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
                [0.5, 0.5, 0.866]])


def _case(shape, nm, nq, seed):
    rng = np.random.default_rng(seed)
    ref = dict(tri=TRI, prism=PRISM, tet=TET, pyr=PYR)[shape]
    verts = np.stack([ref + rng.uniform(-0.1, 0.1, ref.shape) for _ in range(2)])
    if shape == "tri":
        b = O.basis_tri(nm, nq)
        mats = [R.tri_element_mats(nm, nq, nq - 1, v) for v in verts]
        geom = lambda d: O.geom_tri(verts, b, d)
    elif shape == "prism":
        b = O.basis_prism(nm, nq)
        mats = [R.prism_element_mats(nm, nq, nq - 1, v) for v in verts]
        geom = lambda d: O.geom_prism(verts, b, d)
    elif shape == "tet":
        b = O.basis_tet(nm, nq)
        mats = [R.tet_element_mats(nm, nq, nq - 1, nq - 1, v) for v in verts]
        geom = lambda d: O.geom_tet(verts, b, d)
    else:
        b = O.basis_pyr(nm, nq)
        mats = [R.pyr_element_mats(nm, nq, nq, nq - 1, v) for v in verts]
        geom = lambda d: O.geom_pyr(verts, b, d)
    return b, mats, geom, verts


def _ipdb(shape, b, E, deformed, jac, df, f):
    if shape == "tri":
        return O.iprodderiv_tri(b, E, deformed, jac, df, f[0], f[1])
    return dict(prism=O.iprodderiv_prism, tet=O.iprodderiv_tet,
                pyr=O.iprodderiv_pyr)[shape](b, E, deformed, jac, df, f)


@pytest.mark.parametrize("shape", ["tri", "prism", "tet", "pyr"])
@pytest.mark.parametrize("nm,nq", [(3, 4), (5, 6)])
def test_iprodderiv_vs_dense(shape, nm, nq):
    b, mats, geom, verts = _case(shape, nm, nq, nm * 7 + nq)
    E, M, Q = len(verts), b["M"], b["Q"]
    dim = 2 if shape == "tri" else 3
    rng = np.random.default_rng(3)
    f = [rng.uniform(-1, 1, E * Q) for _ in range(dim)]
    # prisms and pyramids on perturbed vertices have per-point factors; the regular form
    # is exact only for affine maps (triangles, tetrahedra)
    forms = [True] + ([False] if shape in ("tri", "tet") else [])
    for deformed in forms:
        jac, df = geom(deformed)
        got = _ipdb(shape, b, E, deformed, jac, df, f)
        for e, m in enumerate(mats):
            WJ = m["W"] * m["J"]
            ref = sum(m["Gphys"][c].T @ (WJ * f[c][e * Q:(e + 1) * Q]) for c in range(dim))
            assert np.linalg.norm(got[e * M:(e + 1) * M] - ref) < 1e-11 * np.linalg.norm(ref), \
                (shape, deformed, e)


@pytest.mark.parametrize("shape", ["tri", "prism", "tet", "pyr"])
def test_iprodderiv_is_the_transpose_of_physderiv(shape):
    nm, nq = 4, 5
    b, _, geom, verts = _case(shape, nm, nq, 11)
    E, M, Q = len(verts), b["M"], b["Q"]
    dim = 2 if shape == "tri" else 3
    jac, df = geom(True)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, E * M)
    f = [rng.uniform(-1, 1, E * Q) for _ in range(dim)]
    bw = getattr(O, f"bwdtrans_{shape}")(b, E, u)
    d = getattr(O, f"physderiv_{shape}")(b, E, True, df, bw)
    w = dict(tri=lambda: np.outer(b["w1"], b["w0"]).ravel(),
             prism=lambda: np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel(),
             tet=lambda: np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel(),
             pyr=lambda: np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel())[shape]()
    WJ = np.tile(w, E) * jac
    rhs = sum(np.dot(f[c], WJ * d[c]) for c in range(dim))
    lhs = np.dot(_ipdb(shape, b, E, True, jac, df, f), u)
    assert abs(lhs - rhs) < 1e-12 * max(abs(rhs), 1.0)
