"""The reference's MatrixFree unit tests for hexahedra and quadrilaterals on the GPU:
UnitTests/Collections/TestHexCollection.cpp (BwdTrans :673, IProductWRTBase :802, :881,
:960, PhysDeriv :2002, :2081, IProductWRTDerivBase :2662, :2752, :2842) and
TestQuadCollection.cpp (BwdTrans :316, IProductWRTBase :957, :1025, :1093, PhysDeriv
:1286, :1353, :1421, IProductWRTDerivBase :1889, :1969, :2148, Helmholtz :2612, :2699):
their elements, orders, points, fields and lambdas.  The reference compares the
MatrixFree operator with the element's own routine to BOOST_CHECK_CLOSE 1e-8 (percent);
here the GPU operator is compared with independent dense matrices (tests/dense_ref.py;
the oracle's collocation derivative for PhysDeriv) to 1e-10 relative.  The hexahedral
Helmholtz cases (:3649, :3747) are in tests/test_gpu_operators.py; the quadrilateral
embedded in three dimensions (:2049, a 2D element with 3D coordinates) is the last test
below."""
import numpy as np
import pytest
import torch

import dense_ref as DR
import ithaca_hip as H
import oracle as O

pytestmark = pytest.mark.gpu

HEX = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1),
       (-1, 1, 1)]
HEX_D = HEX[:6] + [(2, 3, 4)] + HEX[7:]
QUAD = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
QUAD_D = [(-1, -1), (1, -1), (1, 2), (-1, 1)]
QUAD_P = [(-1, -2), (1, -1), (1, 1), (-1, 1)]
QUAD_I = [(-1, -1), (3, -1), (1, 1), (-1, 1)]
QUAD_H = [(-1.5, -1.5), (1, -1), (1, 1), (-1, 1)]

CASES = [  # name, operator, shape, vertices, nm, nq, lambda, elements, affine
    ("TestHexBwdTrans_MatrixFree_UniformP", "BwdTrans", "hex", HEX, 4, 6, 0, 1, True),
    ("TestHexIProductWRTBase_MatrixFree_UniformP_Undeformed", "IProduct", "hex", HEX, 4, 5, 0, 1, True),
    ("TestHexIProductWRTBase_MatrixFree_UniformP_Deformed", "IProduct", "hex", HEX_D, 4, 5, 0, 1, False),
    ("TestHexIProductWRTBase_MatrixFree_UniformP_Deformed_OverInt", "IProduct", "hex", HEX_D, 4, 8, 0, 1, False),
    ("TestHexPhysDeriv_MatrixFree_UniformP_Undeformed", "PhysDeriv", "hex", HEX, 2, 6, 0, 1, True),
    ("TestHexPhysDeriv_MatrixFree_UniformP_Deformed", "PhysDeriv", "hex", HEX_D, 2, 5, 0, 1, False),
    ("TestHexIProductWRTDerivBase_MatrixFree_UniformP_Undeformed", "IProductWRTDerivBase", "hex", HEX, 4, 5, 0, 1, True),
    ("TestHexIProductWRTDerivBase_MatrixFree_UniformP_Deformed", "IProductWRTDerivBase", "hex", HEX_D, 4, 5, 0, 1, False),
    ("TestHexIProductWRTDerivBase_MatrixFree_UniformP_Deformed_OverInt", "IProductWRTDerivBase", "hex", HEX_D, 4, 8, 0, 1, False),
    ("TestQuadBwdTrans_MatrixFree_UniformP", "BwdTrans", "quad", QUAD, 4, 6, 0, 1, True),
    ("TestQuadIProductWRTBase_MatrixFree_UniformP_Undeformed", "IProduct", "quad", QUAD, 5, 6, 0, 1, True),
    ("TestQuadIProductWRTBase_MatrixFree_UniformP_Deformed", "IProduct", "quad", QUAD_D, 5, 6, 0, 1, False),
    ("TestQuadIProductWRTBase_MatrixFree_UniformP_Deformed_OverInt", "IProduct", "quad", QUAD_D, 5, 10, 0, 1, False),
    ("TestQuadPhysDeriv_MatrixFree_UniformP_Undeformed", "PhysDeriv", "quad", QUAD, 2, 5, 0, 1, True),
    ("TestQuadPhysDeriv_MatrixFree_UniformP_Deformed", "PhysDeriv", "quad", QUAD_P, 2, 4, 0, 1, False),
    ("TestQuadIProductWRTDerivBase_MatrixFree_UniformP_Undeformed", "IProductWRTDerivBase", "quad", QUAD, 5, 6, 0, 1, True),
    ("TestQuadIProductWRTDerivBase_MatrixFree_UniformP_Deformed", "IProductWRTDerivBase", "quad", QUAD_I, 5, 6, 0, 1, False),
    ("TestQuadIProductWRTDerivBase_MatrixFree_UniformP_Deformed_OverInt", "IProductWRTDerivBase", "quad", QUAD_I, 5, 10, 0, 1, False),
    ("TestQuadHelmholtz_MatrixFree_UniformP", "Helmholtz", "quad", QUAD, 5, 6, 1.5, 10, True),
    ("TestQuadHelmholtz_MatrixFree_UniformP_Deformed", "Helmholtz", "quad", QUAD_H, 5, 6, 0.0, 10, False),
]


def _fields(X):
    """The tests' fields at the element's points (X: Q x dim)."""
    if X.shape[1] == 2:
        x, y = X[:, 0], X[:, 1]
        return [np.sin(x) * np.cos(y), np.cos(x) * np.sin(y)]
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return [np.sin(x) * np.cos(y) * np.sin(z), np.cos(x) * np.sin(y) * np.cos(z),
            np.cos(x) * np.sin(y) * np.sin(z)]


def _phys_grads(m):
    """d_c = sum_a (A^-1)[a, c] G_a: the basis' physical gradients at the points."""
    dim = len(m["G"])
    return [sum(m["Ainv"][:, a, c][:, None] * m["G"][a] for a in range(dim)) for c in range(dim)]


@pytest.mark.parametrize("name,opname,shape,verts,nm,nq,lam,E,affine", CASES,
                         ids=[c[0] for c in CASES])
def test_reference_tensor_matrixfree_cases(name, opname, shape, verts, nm, nq, lam, E, affine):
    print(f"[case] {name}: nm={nm} nq={nq} E={E}", flush=True)
    v = np.asarray(verts, dtype=np.float64)
    S = shape.capitalize()
    geo = "Regular" if affine or opname == "BwdTrans" else "Deformed"
    b = H.basis_modified_a(nm, nq)
    op = H.Operator(f"{opname}_{S}_{geo}", nm, nq, E, basis=b)
    vs = np.array(np.broadcast_to(v, (E,) + v.shape))
    if opname != "BwdTrans":
        jac, df = H.geom_device(shape, torch.from_numpy(vs).cuda(), nq, b["z"], not affine)
        if op.needs_jac():
            op.set_jac(jac)
        if op.needs_df():
            op.set_df(df)
    m = DR.element_mats(shape, max(nm, 2), nq, v)
    dim = v.shape[1]
    M, Q = m["B"].shape[1], m["B"].shape[0]
    if opname == "PhysDeriv":
        f = _fields(m["X"])[0]
        got = op.apply(torch.from_numpy(np.tile(f, E)).cuda()).cpu().numpy()
        ob = O.basis(nm, nq)
        _, odf = getattr(O, f"geom_{shape}")(vs, nq, ob["z"], not affine)
        ref = O.physderiv(shape, ob, E, not affine, odf, np.tile(f, E))
        for c in range(dim):
            g = got[c * E * Q:(c + 1) * E * Q]
            assert np.linalg.norm(g - ref[c]) < 1e-10 * np.linalg.norm(ref[c]), (name, c)
        return
    WJ = m["W"] * m["J"]
    if opname == "BwdTrans":
        inp, ref = np.ones(M), m["B"] @ np.ones(M)
    elif opname == "IProduct":
        inp = _fields(m["X"])[0]
        ref = m["B"].T @ (WJ * inp)
    elif opname == "IProductWRTDerivBase":
        f = _fields(m["X"])
        d = _phys_grads(m)
        ref = sum(d[c].T @ (WJ * f[c]) for c in range(dim))
        inp = np.concatenate([np.tile(fc, E) for fc in f])
        got = op.apply(torch.from_numpy(inp).cuda()).cpu().numpy().reshape(E, -1)
        for e in range(E):
            assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), (name, e)
        return
    else:
        op.set_lambda(lam)
        inp, ref = np.ones(M), DR.helmholtz_matrix(m, lam) @ np.ones(M)
    got = op.apply(torch.from_numpy(np.tile(inp, E)).cuda()).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), (name, e)


def test_reference_quad_iprodderiv_in_3d_space_case():
    """TestQuadCollection.cpp:2049 TestQuadIProductWRTDerivBase_MatrixFree_UniformP_Deformed_
    ThreeD: the quadrilateral (-1, -1, 0), (3, -1, 0), (1, 1, 1), (-1, 1, 1) in 3D space (a
    trapezoid: per-point factors), 5 modes, 6 GLL points, fields sin x cos y, cos x sin y and
    cos x sin z, the three directions' results summed, 10 elements.  Per point the factors are
    the pseudo-inverse of the 3 x 2 map (dxi_a/dx_c) and J = sqrt(det(A^T A))."""
    print("[case] quad in 3D space, IProductWRTDerivBase", flush=True)
    v = np.array([(-1, -1, 0), (3, -1, 0), (1, 1, 1), (-1, 1, 1)], dtype=np.float64)
    nm, nq, E = 5, 6, 10
    z, w = DR.gll(nq)
    B1, dB1 = DR.modified_a(nm, z)
    G = [np.kron(B1, dB1), np.kron(dB1, B1)]
    W = np.kron(w, w)
    Q = nq * nq
    X, J, P = np.zeros((Q, 3)), np.zeros(Q), np.zeros((Q, 2, 3))
    for j in range(nq):
        for i in range(nq):
            x, A = DR.quad_map(v, np.array([z[i], z[j]]))
            X[j * nq + i], P[j * nq + i] = x, np.linalg.pinv(A)
            J[j * nq + i] = np.sqrt(np.linalg.det(A.T @ A))
    assert np.ptp(J) > 1e-3  # not a parallelogram: the deformed path
    f = [np.sin(X[:, 0]) * np.cos(X[:, 1]), np.cos(X[:, 0]) * np.sin(X[:, 1]),
         np.cos(X[:, 0]) * np.sin(X[:, 2])]
    ref = sum((P[:, a, c][:, None] * G[a]).T @ (W * J * f[c]) for c in range(3) for a in range(2))
    op = H.Operator("IProductWRTDerivBase_Quad_Deformed", nm, nq, E, coord_dim=3)
    assert op.in_size == 3 * E * Q
    op.set_jac(np.tile(J, E))
    op.set_df(np.concatenate([np.tile(P[:, a, c], E) for c in range(3) for a in range(2)]))
    inp = torch.from_numpy(np.concatenate([np.tile(fc, E) for fc in f])).cuda()
    got = op.apply(inp).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), e
    # the same operator on two coordinates is unchanged by the setter's default
    op2 = H.Operator("IProductWRTDerivBase_Quad_Deformed", nm, nq, E)
    assert op2.in_size == 2 * E * Q
    with pytest.raises(H.MFError):
        H.Operator("IProductWRTDerivBase_Quad_Deformed", nm, nq, E, precision=H.FP32,
                   coord_dim=3)
