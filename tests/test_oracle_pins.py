"""Pins of the CPU oracle: reference golden values, closed forms, dense assembly.

Golden values come from the reference's own regression specifications
(library/Demos/LocalRegions/Tests/*.tst, library/Demos/StdRegions/Tests/*.tst); the
LocProject demo (library/Demos/LocalRegions/LocProject.cpp:380-495) evaluates
sum_{k,l,m<6} x^k y^l z^m at the quadrature points, projects it (IProduct + mass solve),
back-transforms and prints the L2 / Linf errors.
"""
import numpy as np
import pytest

import dense_ref as R
import oracle as O

DEF_HEX = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.5, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0,
           1.5, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.5]
REG_HEX = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0,
           1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0]
DEF_QUAD = [0.0, 0.0, 1.0, 0.0, 1.5, 1.5, 0.0, 1.0]


def _poly_sol(X, dim, diff):
    x, y = X[0], X[1]
    z = X[2] if dim == 3 else np.zeros_like(x)
    pw = lambda a, i: np.ones_like(a) if i < 0 else a ** i
    sol = np.zeros_like(x)
    for k in range(6):
        for l in range(6):
            for m in range(6 if dim == 3 else 1):
                if not diff:
                    sol += pw(x, k) * pw(y, l) * pw(z, m)
                else:
                    sol += k * pw(x, k - 1) * pw(y, l) * pw(z, m)
                    sol += l * pw(x, k) * pw(y, l - 1) * pw(z, m)
                    sol += m * pw(x, k) * pw(y, l) * pw(z, m - 1)
    return sol


def _locproject(shape, coords, diff):
    dim = 3 if shape == "hex" else 2
    b = O.basis(6, 7)
    v = np.array(coords).reshape(1, 2 ** dim, dim)
    geom = O.geom_hex if dim == 3 else O.geom_quad
    crd = O.coords_hex if dim == 3 else O.coords_quad
    jac, df = geom(v, 7, b["z"], True)
    X = crd(v, 7, b["z"])
    sol = _poly_sol(X, dim, False)
    if diff:
        ds = O.physderiv(shape, b, 1, True, df, sol)
        sol = sum(ds)
    M = 6 ** dim
    mass = np.zeros((M, M))
    for i in range(M):
        e = np.zeros(M); e[i] = 1.0
        mass[:, i] = O.iproduct(shape, b, 1, True, jac, O.bwdtrans(shape, b, 1, e))
    coef = np.linalg.solve(mass, O.iproduct(shape, b, 1, True, jac, sol))
    phys = O.bwdtrans(shape, b, 1, coef)
    if diff:
        sol = _poly_sol(X, dim, True)
    W = b["w"]
    Wq = np.kron(np.kron(W, W), W) if dim == 3 else np.kron(W, W)
    l2 = np.sqrt(np.sum(Wq * jac * (phys - sol) ** 2))
    return l2, np.abs(phys - sol).max()


@pytest.mark.parametrize("shape,coords,diff,l2,linf", [
    # LocProject3D_Def_Hex_Mod_Basis_P6_Q7.tst
    ("hex", DEF_HEX, False, 0.00820104, 0.0804162),
    # LocProject_Diff3D_Lin_Deformed_Hex_Mod_Basis_P6_Q7.tst
    ("hex", DEF_HEX, True, 0.0353098, 0.446618),
    # LocProject2D_Def_Quad_Mod_Basis_P6_Q7.tst
    ("quad", DEF_QUAD, False, 0.0970116, 0.793473),
    # LocProject_Diff2D_Lin_Deformed_Quad_Mod_Basis_P6_Q7.tst
    ("quad", DEF_QUAD, True, 0.288711, 2.40284),
])
def test_locproject_golden(shape, coords, diff, l2, linf):
    got_l2, got_linf = _locproject(shape, coords, diff)
    # goldens are printed with 6 significant digits
    assert abs(got_l2 - l2) <= 6e-6 * l2
    assert abs(got_linf - linf) <= 6e-6 * linf


@pytest.mark.parametrize("diff,l2_max", [(False, 5.65299e-11 + 1e-8), (True, 2.74838e-10 + 1e-9)])
def test_locproject_regular_hex(diff, l2_max):
    # LocProject3D_Hex_Mod_Basis_P6_Q7.tst / LocProject_Diff3D_Reg_Hex_Mod_Basis_P6_Q7.tst
    l2, _ = _locproject("hex", REG_HEX, diff)
    assert l2 <= l2_max


@pytest.mark.parametrize("nq", [2, 3, 5, 6, 7, 9, 12, 16])
def test_gll_closed_forms(nq):
    b = O.basis(2, nq)
    z, w, D = b["z"], b["w"], b["D"].reshape(nq, nq)
    zr, wr = R.gll(nq)
    assert np.allclose(z, zr, atol=1e-14, rtol=0)
    assert np.allclose(w, wr, atol=1e-14, rtol=0)
    assert abs(w.sum() - 2.0) < 1e-14
    assert abs(w[0] - 2.0 / (nq * (nq - 1))) < 1e-15
    # quadrature exact to degree 2nq-3
    for deg in range(2 * nq - 2):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.sum(w * z ** deg) - exact) < 1e-13
    # D[a, b] = l_a'(z_b): derivative of z^k is exact for k < nq
    for k in range(nq):
        dk = D.T @ (z ** k)
        assert np.allclose(dk, k * z ** max(k - 1, 0) if k else 0.0, atol=1e-11)


@pytest.mark.parametrize("nm,nq", [(2, 2), (3, 4), (5, 6), (6, 7), (8, 16)])
def test_modified_a_closed_form(nm, nq):
    b = O.basis(nm, nq)
    B, dB = R.modified_a(nm, b["z"])
    assert np.allclose(b["bdata"], B.T.ravel(), atol=1e-14, rtol=0)
    assert np.allclose(b["dbdata"], dB.T.ravel(), atol=1e-11, rtol=0)


def test_geometry_closed_forms():
    b = O.basis(6, 7)
    h = 0.25
    v = np.array(REG_HEX).reshape(1, 8, 3) * h + np.array([0.5, 0.25, 0.0])
    jac, df = O.geom_hex(v, 7, b["z"], deformed=False)
    assert abs(jac[0] - (h / 2) ** 3) < 1e-16
    assert np.allclose(df.reshape(9), np.eye(3).ravel() * 2 / h, atol=1e-13)
    rng = np.random.default_rng(7)
    v = np.array(REG_HEX).reshape(8, 3) + rng.uniform(-0.2, 0.2, (8, 3))
    jac, df = O.geom_hex(v[None], 5, O.basis(2, 5)["z"], True)
    z5 = O.basis(2, 5)["z"]
    df = df.reshape(9, -1)
    for q in [0, 17, 124]:
        i, j, k = q % 5, (q // 5) % 5, q // 25
        _, A = R.hex_map(v, np.array([z5[i], z5[j], z5[k]]))
        assert abs(jac[q] - abs(np.linalg.det(A))) < 1e-14
        # df[c*3+a] = dxi_a/dx_c = (A^{-1})[a, c]
        Ainv = np.linalg.inv(A)
        assert np.allclose(df[:, q].reshape(3, 3), Ainv.T, atol=1e-13)


def _rand_hex(rng, n=3):
    base = np.array(REG_HEX).reshape(8, 3)
    return np.stack([base + rng.uniform(-0.15, 0.15, (8, 3)) for _ in range(n)])


def _rand_quad(rng, n=3):
    base = np.array([0.0, 0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0]).reshape(4, 2)
    return np.stack([base + rng.uniform(-0.15, 0.15, (4, 2)) for _ in range(n)])


@pytest.mark.parametrize("shape,nm,nq", [("hex", 6, 7), ("hex", 3, 5), ("hex", 4, 8),
                                         ("quad", 5, 6), ("quad", 6, 7), ("quad", 3, 6)])
def test_operators_vs_dense(shape, nm, nq):
    rng = np.random.default_rng(11)
    verts = _rand_hex(rng) if shape == "hex" else _rand_quad(rng)
    E = verts.shape[0]
    b = O.basis(nm, nq)
    geom = O.geom_hex if shape == "hex" else O.geom_quad
    jac, df = geom(verts, nq, b["z"], True)
    dim = 3 if shape == "hex" else 2
    M, Q = nm ** dim, nq ** dim
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    bw = O.bwdtrans(shape, b, E, u)
    ip = O.iproduct(shape, b, E, True, jac, f)
    hh = O.helmholtz(shape, b, E, True, jac, df, 3.5, u)
    for e in range(E):
        m = R.element_mats(shape, nm, nq, verts[e])
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        ref_bw = m["B"] @ ue
        ref_ip = m["B"].T @ (m["W"] * m["J"] * fe)
        ref_h = R.helmholtz_matrix(m, 3.5) @ ue
        rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
        assert rel(bw[e * Q:(e + 1) * Q], ref_bw) < 1e-13
        assert rel(ip[e * M:(e + 1) * M], ref_ip) < 1e-13
        assert rel(hh[e * M:(e + 1) * M], ref_h) < 1e-12


@pytest.mark.parametrize("shape,nm,nq", [("hex", 6, 7), ("hex", 3, 5), ("quad", 5, 6),
                                         ("quad", 4, 8)])
def test_deriv_operators_vs_dense(shape, nm, nq):
    """PhysDeriv of a field of the space equals its exact physical gradient, and
    IProductWRTDerivBase equals sum_c (d_c B)^T diag(W J) in_c, both with d_c B built from
    independent numpy code (closed-form basis derivatives and the inverse map Jacobian)."""
    rng = np.random.default_rng(13)
    verts = _rand_hex(rng) if shape == "hex" else _rand_quad(rng)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    b = O.basis(nm, nq)
    geom = O.geom_hex if shape == "hex" else O.geom_quad
    jac, df = geom(verts, nq, b["z"], True)
    M, Q = nm ** dim, nq ** dim
    u = rng.uniform(-1, 1, E * M)
    ins = [rng.uniform(-1, 1, E * Q) for _ in range(dim)]
    grad = O.physderiv(shape, b, E, True, df, O.bwdtrans(shape, b, E, u))
    ipd = O.iprod_deriv(shape, b, E, True, jac, df, ins)
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    for e in range(E):
        m = R.element_mats(shape, nm, nq, verts[e])
        dc = [sum(m["Ainv"][:, a, c][:, None] * m["G"][a] for a in range(dim))
              for c in range(dim)]
        ue = u[e * M:(e + 1) * M]
        sl = slice(e * Q, (e + 1) * Q)
        for c in range(dim):
            assert rel(grad[c][sl], dc[c] @ ue) < 1e-12
        WJ = m["W"] * m["J"]
        ref = sum(dc[c].T @ (WJ * ins[c][sl]) for c in range(dim))
        assert rel(ipd[e * M:(e + 1) * M], ref) < 1e-12


def test_regular_equals_deformed_on_affine_elements():
    # a parallelepiped is "regular": per-element factors must give the same operator
    b = O.basis(6, 7)
    A = np.array([[1.0, 0.2, 0.1], [0.0, 0.8, 0.3], [0.1, 0.0, 1.2]])
    v = (np.array(REG_HEX).reshape(8, 3) @ A.T)[None]
    jd, dd = O.geom_hex(v, 7, b["z"], True)
    jr, dr = O.geom_hex(v, 7, b["z"], False)
    u = np.random.default_rng(3).uniform(-1, 1, 216)
    hd = O.helmholtz("hex", b, 1, True, jd, dd, 2.0, u)
    hr = O.helmholtz("hex", b, 1, False, jr, dr, 2.0, u)
    assert np.linalg.norm(hd - hr) / np.linalg.norm(hd) < 1e-13


def test_helmholtz_invariants():
    rng = np.random.default_rng(5)
    verts = _rand_hex(rng, 4)
    b = O.basis(6, 7)
    jac, df = O.geom_hex(verts, 7, b["z"], True)
    E, M = 4, 216
    # constants: vertex modes all one (phi_0 + phi_1 = 1 per direction)
    one = np.zeros(M)
    for r in range(2):
        for q in range(2):
            for p in range(2):
                one[(r * 6 + q) * 6 + p] = 1.0
    ones = np.tile(one, E)
    lap1 = O.helmholtz("hex", b, E, True, jac, df, 0.0, ones)
    assert np.abs(lap1).max() < 1e-12
    # IProduct(1) . 1 = element volumes; mass term = IProduct o BwdTrans
    mass1 = O.helmholtz("hex", b, E, True, jac, df, 1.0, ones)
    ip1 = O.iproduct("hex", b, E, True, jac, np.ones(E * 343))
    assert np.allclose(mass1, ip1, atol=1e-13)
    for e in range(E):
        assert abs(ip1[e * M:(e + 1) * M] @ one - np.sum(b["w"][:, None, None] * b["w"][None, :, None]
                   * b["w"][None, None, :] * jac[e * 343:(e + 1) * 343].reshape(7, 7, 7))) < 1e-13
    # symmetry and positive definiteness for lambda > 0
    u, v = rng.uniform(-1, 1, E * M), rng.uniform(-1, 1, E * M)
    Hu = O.helmholtz("hex", b, E, True, jac, df, 10.0, u)
    Hv = O.helmholtz("hex", b, E, True, jac, df, 10.0, v)
    assert abs(v @ Hu - u @ Hv) < 1e-12 * abs(v @ Hu)
    assert u @ Hu > 0


@pytest.mark.parametrize("E,deformed", [(1, True), (6, True), (7, False)])
def test_avx2_helmholtz_vs_dense(E, deformed):
    """The width-4 (AVX2 SIMD ABI) restatement against the dense operator, with element
    counts that leave a partly filled last vector."""
    rng = np.random.default_rng(17)
    verts = _rand_hex(rng, E)
    if not deformed:
        verts = np.stack([np.array(REG_HEX).reshape(8, 3) * (1 + 0.1 * e) + e for e in range(E)])
    b = O.basis(5, 7)
    jac, df = O.geom_hex(verts, 7, b["z"], deformed)
    u = rng.uniform(-1, 1, E * 125)
    h4 = O.helmholtz("hex", b, E, deformed, jac, df, 2.5, u, simd="avx2")
    h1 = O.helmholtz("hex", b, E, deformed, jac, df, 2.5, u)
    assert np.abs(h4 - h1).max() <= 1e-13 * np.abs(h1).max()
    for e in range(E):
        m = R.element_mats("hex", 5, 7, verts[e])
        ref = R.helmholtz_matrix(m, 2.5) @ u[e * 125:(e + 1) * 125]
        assert np.linalg.norm(h4[e * 125:(e + 1) * 125] - ref) < 1e-12 * np.linalg.norm(ref)
