"""The reference's MatrixFree IProductWRTDerivBase unit tests for the collapsed shapes, on
the GPU: the elements, orders, point sets and fields of
UnitTests/Collections/TestTriCollection.cpp:1956-, :2133- (TestTriIProductWRTDerivBase_
MatrixFree_UniformP_{Undeformed,Deformed}), TestTetCollection.cpp:2537-,
TestPrismCollection.cpp (TestPrismIProductWRTDerivBase_MatriFree_UniformP_{Undeformed,
Deformed,Deformed_OverInt}_MultiElmt) and TestPyrCollection.cpp:2084-
(TestPyrIProductWRTDerivBase_MatrixFree_Deformed_MultiElmt).  Those tests compare the
MatrixFree operator with the element's own IProductWRTDerivBase summed over directions to
BOOST_CHECK_CLOSE 1e-8 (percent); here the GPU operator is compared with the same sum
formed from independent dense gradient matrices (tests/dense_ref.py), to 1e-10
relative, the element repeated over several elements as in the MultiElmt cases."""
import numpy as np
import pytest
import torch

import dense_ref as R
import ithaca_hip as H

pytestmark = pytest.mark.gpu

E = 10

CASES = [
    # name, shape, vertices, nm, (nq0, nq1[, nq2]), affine
    ("tri_undeformed", "tri", [(-1, -1), (1, -1), (-1, 1)], 4, (5, 4), True),
    ("tri_deformed", "tri", [(-1, -1), (1, -1), (-1, 2)], 4, (5, 4), True),
    ("tet_undeformed", "tet", [(-2, -3, -4), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)], 4,
     (5, 4, 4), True),
    ("prism_undeformed", "prism", [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                   (-1, -1, 1), (-1, 1, 1)], 6, (7, 7, 6), True),
    ("prism_deformed", "prism", [(-2, -3, -4), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                 (-1, -1, 1), (-1, 1, 1)], 6, (7, 7, 6), False),
    ("prism_deformed_overint", "prism", [(-2, -3, -4), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                         (-1, -1, 1), (-1, 1, 1)], 6, (12, 12, 11), False),
    ("pyr_deformed", "pyr", [(-1.5, -1.5, -1.5), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                             (-1, -1, 1)], 4, (5, 5, 4), False),
]


def _mats(shape, nm, nq, verts):
    if shape == "tri":
        return R.tri_element_mats(nm, nq[0], nq[1], verts)
    if shape == "tet":
        return R.tet_element_mats(nm, nq[0], nq[1], nq[2], verts)
    if shape == "prism":
        return R.prism_element_mats(nm, nq[0], nq[2], verts)
    return R.pyr_element_mats(nm, nq[0], nq[1], nq[2], verts)


def _fields(X):
    """The unit tests' fields at the element's points (X: dim x Q coordinates)."""
    if X.shape[0] == 2:
        x, y = X
        return [np.sin(x) * np.cos(y), np.cos(x) * np.sin(y)]
    x, y, z = X
    return [np.sin(x) * np.cos(y) * np.sin(z), np.cos(x) * np.sin(y) * np.cos(z),
            np.cos(x) * np.sin(y) * np.sin(z)]


@pytest.mark.parametrize("name,shape,verts,nm,nq,affine", CASES, ids=[c[0] for c in CASES])
def test_reference_iproduct_wrt_deriv_base_cases(name, shape, verts, nm, nq, affine):
    print(f"[case] {name}: nm={nm} nq={nq}", flush=True)
    v = np.asarray(verts, dtype=np.float64)
    m = _mats(shape, nm, nq, v)
    dim = v.shape[1]
    f = _fields(m["X"])
    WJ = m["W"] * m["J"]
    ref = sum(m["Gphys"][c].T @ (WJ * f[c]) for c in range(dim))
    key = f"IProductWRTDerivBase_{shape.capitalize()}_{'Regular' if affine else 'Deformed'}"
    op = H.Operator(key, nm, nq, E)
    vd = torch.from_numpy(np.array(np.broadcast_to(v, (E,) + v.shape))).cuda()
    jac, df = H.geom_device(shape, vd, nq, None, not affine)
    op.set_jac(jac)
    op.set_df(df)
    inp = torch.from_numpy(np.concatenate([np.tile(fc, E) for fc in f])).cuda()
    got = op.apply(inp).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), (name, e)


# The other MatrixFree cases of the same files: BwdTrans of unit coefficients, IProduct of
# sin x cos y (sin z), PhysDeriv of that field, Helmholtz of unit coefficients; the
# reference compares with the element's BwdTrans, IProductWRTBase, PhysDeriv and
# GeneralMatrixOp(eHelmholtz).  (name, operator, shape, vertices, nm, numQuadPoints,
# lambda, elements); the point counts per direction follow the tests' PointsKeys.
TRI, TRI_D = [(-1, -1), (1, -1), (-1, 1)], [(-1, -1), (1, -1), (-1, 2)]
TRI_H = [(-1.5, -1.5), (1, -1), (-1, 1)]
TET = [(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]
TET_D = [(-2, -3, -4), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]
PRI = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (-1, 1, 1)]
PRI_D = [(-2, -3, -4)] + PRI[1:]
PYR = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1)]
PYR_D, PYR_H = [(-2, -3, -4)] + PYR[1:], [(-1.5, -1.5, -1.5)] + PYR[1:]
OTHER = [
    ("TestTriBwdTrans_MatrixFree_UniformP", "BwdTrans", "tri", TRI, 4, 5, 0, 1),
    ("TestTriBwdTrans_MatrixFree_UniformP_OverInt", "BwdTrans", "tri", TRI, 4, 8, 0, 1),
    ("TestTriIProductWRTBase_MatrixFree_UniformP_Undeformed", "IProduct", "tri", TRI, 4, 5, 0, 1),
    ("TestTriIProductWRTBase_MatrixFree_UniformP_Deformed", "IProduct", "tri", TRI_D, 4, 5, 0, 1),
    ("TestTriIProductWRTBase_MatrixFree_UniformP_Deformed_OverInt", "IProduct", "tri", TRI_D, 4, 8, 0, 1),
    ("TestTriPhysDeriv_MatrixFree_UniformP_Undeformed", "PhysDeriv", "tri", TRI, 2, 5, 0, 1),
    ("TestTriPhysDeriv_MatrixFree_UniformP_Deformed", "PhysDeriv", "tri", TRI_D, 2, 5, 0, 1),
    ("TestTriHelmholtz_MatrixFree_UniformP", "Helmholtz", "tri", TRI_H, 4, 5, 1.5, 10),
    ("TestTriHelmholtz_MatrixFree_UniformP_OverInt", "Helmholtz", "tri", TRI_H, 4, 8, 1.5, 10),
    ("TestTetBwdTrans_MatrixFree_UniformP", "BwdTrans", "tet", TET, 4, 5, 0, 1),
    ("TestTetBwdTrans_MatrixFree_UniformP_OverInt", "BwdTrans", "tet", TET, 4, 8, 0, 1),
    ("TestTetIProductWRTBase_MatrixFree_UniformP_Undeformed", "IProduct", "tet", TET, 4, 5, 0, 1),
    ("TestTetIProductWRTBase_MatrixFree_UniformP_Deformed", "IProduct", "tet", TET_D, 4, 5, 0, 1),
    ("TestTetIProductWRTBase_MatrixFree_UniformP_Deformed_OverInt", "IProduct", "tet", TET_D, 4, 8, 0, 1),
    ("TestTetPhysDeriv_MatrixFree_UniformP", "PhysDeriv", "tet", TET, 4, 5, 0, 1),
    ("TestTetmHelmholtz_MatrixFree_UniformP", "Helmholtz", "tet", TET, 4, 5, 0.0, 10),
    ("TestTetmHelmholtz_MatrixFree_Deformed_OverInt", "Helmholtz", "tet", TET_D, 4, 8, 0.0, 10),
    ("TestPrismBwdTrans_MatrixFree_UniformP_MultiElmt", "BwdTrans", "prism", PRI, 4, 5, 0, 2),
    ("TestPrismBwdTrans_MatrixFree_UniformP_OverInt_MultiElmt", "BwdTrans", "prism", PRI, 4, 8, 0, 2),
    ("TestPrismIProductWRTBase_MatrixFree_UniformP_Undeformed_MultiElmt", "IProduct", "prism", PRI, 4, 5, 0, 2),
    ("TestPrismIProductWRTBase_MatrixFree_UniformP_Deformed_MultiElmt", "IProduct", "prism", PRI_D, 4, 5, 0, 2),
    ("TestPrismIProductWRTBase_MatrixFree_UniformP_Deformed_OverInt_MultiElmt", "IProduct", "prism", PRI_D, 4, 8, 0, 2),
    ("TestPrismPhysDeriv_MatrixFree_UniformP_MultiElmt", "PhysDeriv", "prism", PRI, 2, 5, 0, 2),
    ("TestPrismHelmholtz_MatrixFree_UniformP", "Helmholtz", "prism", PRI, 6, 7, 0.0, 10),
    ("TestPrismHelmholtz_MatrixFree_Deformed_OverInt", "Helmholtz", "prism", PRI_D, 6, 10, 0.0, 10),
    ("TestPyrPhysDeriv_MatrixFree_UniformP_MultiElmt", "PhysDeriv", "pyr", PYR, 2, 5, 0, 10),
    ("TestPyrIProductWRTBase_MatrixFree_UniformP_MultiElmt", "IProduct", "pyr", PYR, 4, 5, 0, 10),
    ("TestPyrIProductWRTBase_MatrixFree_Deformed_MultiElmt", "IProduct", "pyr", PYR_D, 4, 5, 0, 5),
    ("TestPyrBwdTrans_MatrixFree_UniformP_MultiElmt", "BwdTrans", "pyr", PYR, 4, 5, 0, 10),
    ("TestPyrBwdTrans_MatrixFree_UniformP_OverInt_MultiElmt", "BwdTrans", "pyr", PYR, 4, 8, 0, 10),
    ("TestPyrHelmholtz_MatrixFree_UniformP", "Helmholtz", "pyr", PYR, 4, 5, 1.5, 10),
    ("TestPyrHelmholtz_MatrixFree_Deformed_overInt", "Helmholtz", "pyr", PYR_H, 4, 8, 1.5, 10),
]


def _points(shape, nq):
    """Points per direction of the tests' PointsKeys (GLL, then Gauss-Radau with one fewer)."""
    return {"tri": (nq, nq - 1), "tet": (nq, nq - 1, nq - 1), "prism": (nq, nq, nq - 1),
            "pyr": (nq, nq, nq - 1)}[shape]


def _affine(shape, v):
    """Straight-sided triangles and tetrahedra are affine; prisms and pyramids when their
    quadrilateral base is a parallelogram (v0 + v2 = v1 + v3), as for the reference
    elements used here."""
    if shape in ("tri", "tet"):
        return True
    return bool(np.allclose(v[0] + v[2], v[1] + v[3]))


def _oracle_physderiv(shape, nm, nq, v, E, affine, f):
    import oracle as O
    verts = np.ascontiguousarray(np.broadcast_to(v, (E,) + v.shape))
    if shape == "tri":
        b = O.basis_tri(nm, nq[0], nq[1])
    elif shape == "tet":
        b = O.basis_tet(nm, nq[0], nq[1], nq[2])
    elif shape == "prism":
        b = O.basis_prism(nm, nq[0], nq[2])
    else:
        b = O.basis_pyr(nm, nq[0], nq[2])
    _, df = getattr(O, f"geom_{shape}")(verts, b, not affine)
    return getattr(O, f"physderiv_{shape}")(b, E, not affine, df, np.tile(f, E))


@pytest.mark.parametrize("name,opname,shape,verts,nm,nqb,lam,E", OTHER, ids=[c[0] for c in OTHER])
def test_reference_matrixfree_cases(name, opname, shape, verts, nm, nqb, lam, E):
    nq = _points(shape, nqb)
    print(f"[case] {name}: nm={nm} nq={nq} E={E}", flush=True)
    v = np.asarray(verts, dtype=np.float64)
    affine = _affine(shape, v)
    geo = "Regular" if affine or opname == "BwdTrans" else "Deformed"
    op = H.Operator(f"{opname}_{shape.capitalize()}_{geo}", nm, nq, E)
    vd = torch.from_numpy(np.array(np.broadcast_to(v, (E,) + v.shape))).cuda()
    if opname != "BwdTrans":
        jac, df = H.geom_device(shape, vd, nq, None, not affine)
        if op.needs_jac():
            op.set_jac(jac)
        if op.needs_df():
            op.set_df(df)
    if opname == "PhysDeriv":
        x = _fields(np.stack(_coords(shape, nm, nq, v)))[0]
        got = op.apply(torch.from_numpy(np.tile(x, E)).cuda()).cpu().numpy()
        ref = _oracle_physderiv(shape, nm, nq, v, E, affine, x)
        Q = x.size
        for c in range(len(ref)):
            g, r = got[c * E * Q:(c + 1) * E * Q], ref[c]
            assert np.linalg.norm(g - r) < 1e-10 * np.linalg.norm(r), (name, c)
        return
    m = _mats(shape, nm, nq, v)
    M, Q = m["B"].shape[1], m["B"].shape[0]
    if opname == "BwdTrans":
        inp, ref = np.ones(M), m["B"] @ np.ones(M)
    elif opname == "IProduct":
        inp = _fields(m["X"])[0]
        ref = m["B"].T @ (m["W"] * m["J"] * inp)
    else:
        op.set_lambda(lam)
        helm = getattr(R, f"{shape}_helmholtz_matrix")(m, lam)
        inp, ref = np.ones(M), helm @ np.ones(M)
    got = op.apply(torch.from_numpy(np.tile(inp, E)).cuda()).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), (name, e)


def _coords(shape, nm, nq, v):
    """Physical coordinates of the element's points (the dense reference's map)."""
    return list(_mats(shape, max(nm, 2), nq, v)["X"])


@pytest.mark.parametrize("deformed", [True, False])
def test_reference_tri_iprodderiv_in_3d_space_case(deformed):
    """TestTriCollection.cpp:2233 TestTriIProductWRTDerivBase_MatrixFree_UniformP_Deformed_
    MultiElmt_ThreeD: the triangle (-1, -1, 0), (1, -1, 0), (-1, 2, 1) in 3D space, 4 modes,
    5 x 4 points, fields sin x cos y, cos x sin y and cos x sin z, the three directions'
    results summed.  The derivative factors of a surface are the pseudo-inverse of its 3 x 2
    map (dxi_a/dx_c, six per point), J = sqrt(det(A^T A))."""
    print(f"[case] tri in 3D space, IProductWRTDerivBase deformed={deformed}", flush=True)
    v = np.array([(-1, -1, 0), (1, -1, 0), (-1, 2, 1)], dtype=np.float64)
    nm, nq = 4, (5, 4)
    m = R.tri_element_mats(nm, nq[0], nq[1], v)
    x, y, z = m["X"]
    f = [np.sin(x) * np.cos(y), np.cos(x) * np.sin(y), np.cos(x) * np.sin(z)]
    ref = sum(m["Gphys"][c].T @ (m["W"] * m["J"] * f[c]) for c in range(3))
    Q = nq[0] * nq[1]
    key = f"IProductWRTDerivBase_Tri_{'Deformed' if deformed else 'Regular'}"
    op = H.Operator(key, nm, nq, E, coord_dim=3)
    assert op.in_size == 3 * E * Q
    n = E * Q if deformed else E
    op.set_jac(np.full(n, m["J"]))
    # GetDerivFactors order: factor 2c + a is dxi_a / dx_c
    op.set_df(np.concatenate([np.full(n, m["Ainv"][a, c]) for c in range(3) for a in range(2)]))
    inp = torch.from_numpy(np.concatenate([np.tile(fc, E) for fc in f])).cuda()
    got = op.apply(inp).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), e


def test_coord_dim_is_refused_where_the_reference_has_no_3d_form():
    print("[case] coord_dim refusals", flush=True)
    with pytest.raises(H.MFError):
        H.Operator("PhysDeriv_Tri_Regular", 4, (5, 4), E, coord_dim=3)
    op = H.Operator("IProductWRTDerivBase_Tri_Regular", 4, (5, 4), E)
    op.set_jac(np.ones(E))
    op.set_df(np.zeros(4 * E))
    with pytest.raises(H.MFError):  # after SetDF
        H._check(H.lib().imf_set_coord_dim(op._h, 3))
    with pytest.raises(H.MFError):
        H._check(H.lib().imf_set_coord_dim(op._h, 4))
