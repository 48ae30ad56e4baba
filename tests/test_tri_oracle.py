"""Pins of the triangle oracle (collapsed coordinates): Gauss-Radau points against scipy,
Modified_B against closed forms, the four operators against dense matrices from
independent numpy code, and the reference's LocProject triangle goldens
(library/Demos/LocalRegions/Tests/LocProject2D_Tri_Mod_Basis_P6_Q7.tst,
LocProject_Diff2D_Tri_Mod_Basis_P6_Q7.tst)."""
import numpy as np
import pytest

import dense_ref as R
import oracle as O


@pytest.mark.parametrize("nq", [2, 3, 6, 7, 9, 15])
def test_gauss_radau_points_and_weights(nq):
    b = O.basis_tri(3, nq + 1, nq)
    z, w = R.gauss_radau_a1(nq)
    assert np.abs(b["z1"] - z).max() < 1e-14
    assert np.abs(b["w1_raw"] - w).max() < 1e-14
    # exact for (1-z) p(z), deg p <= 2 nq - 2
    for k in range(2 * nq - 1):
        exact = (1 - (-1) ** (k + 1)) / (k + 1) - (1 - (-1) ** (k + 2)) / (k + 2)
        assert abs(np.sum(b["w1_raw"] * b["z1"] ** k) - exact) < 1e-13


@pytest.mark.parametrize("nm,nq", [(2, 2), (4, 5), (6, 6), (8, 15)])
def test_modified_b_closed_form(nm, nq):
    b = O.basis_tri(nm, nq + 1, nq)
    B = R.modified_b(nm, b["z1"])
    assert np.abs(b["bdata1"].reshape(-1, nq).T - B).max() < 1e-13
    # dbdata: exact derivative of the interpolating polynomial of each mode
    for m in range(B.shape[1]):
        c = np.polynomial.polynomial.polyfit(b["z1"], B[:, m], nq - 1)
        d = np.polynomial.polynomial.polyval(b["z1"], np.polynomial.polynomial.polyder(c))
        assert np.abs(b["dbdata1"][m * nq:(m + 1) * nq] - d).max() < 1e-10 * max(1, np.abs(d).max())


TRIS = np.array([[[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]],
                 [[0.2, -0.1], [1.3, 0.4], [0.5, 1.1]],
                 [[0.0, 0.0], [1.0, 1.0], [0.5, 1.0]]])


@pytest.mark.parametrize("nm,nq0", [(6, 7), (3, 5), (4, 8), (2, 3), (8, 16)])
@pytest.mark.parametrize("deformed", [False, True])
def test_tri_operators_vs_dense(nm, nq0, deformed):
    b = O.basis_tri(nm, nq0)
    E = len(TRIS)
    jac, df = O.geom_tri(TRIS, b, deformed)
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(nm * 10 + nq0)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    bw = O.bwdtrans_tri(b, E, u)
    ip = O.iproduct_tri(b, E, deformed, jac, f)
    hh = O.helmholtz_tri(b, E, deformed, jac, df, 2.5, u)
    g0, g1 = O.physderiv_tri(b, E, deformed, df, bw)
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    for e in range(E):
        m = R.tri_element_mats(nm, nq0, nq0 - 1, TRIS[e])
        assert m["fit_residual"] < 1e-12
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        assert rel(bw[e * Q:(e + 1) * Q], m["B"] @ ue) < 1e-13
        assert rel(ip[e * M:(e + 1) * M], m["B"].T @ (m["W"] * m["J"] * fe)) < 1e-13
        assert rel(hh[e * M:(e + 1) * M], R.tri_helmholtz_matrix(m, 2.5) @ ue) < 1e-11
        assert rel(g0[e * Q:(e + 1) * Q], m["Gphys"][0] @ ue) < 1e-11
        assert rel(g1[e * Q:(e + 1) * Q], m["Gphys"][1] @ ue) < 1e-11


def test_tri_helmholtz_invariants():
    b = O.basis_tri(6, 7)
    E = len(TRIS)
    jac, df = O.geom_tri(TRIS, b)
    M = b["M"]
    # the constant field: vertex modes 0, 1 (collapsed top vertex) and the p = 1 vertex
    one = np.zeros(M); one[0] = one[1] = 1.0; one[6] = 1.0
    ones = np.tile(one, E)
    assert np.abs(O.bwdtrans_tri(b, E, ones) - 1.0).max() < 1e-13
    assert np.abs(O.helmholtz_tri(b, E, False, jac, df, 0.0, ones)).max() < 1e-12
    # IProduct(1) . 1 = area
    ip = O.iproduct_tri(b, E, False, jac, np.ones(E * b["Q"]))
    for e in range(E):
        v = TRIS[e]
        d1, d2 = v[1] - v[0], v[2] - v[0]
        area = 0.5 * abs(d1[0] * d2[1] - d1[1] * d2[0])
        assert abs(ip[e * M:(e + 1) * M] @ one - area) < 1e-13


def _tri_sol(x, y, diff):
    s = np.zeros_like(x)
    for k in range(6):
        for l in range(6 - k):
            if not diff:
                s += x ** k * y ** l
            else:
                s += k * x ** max(k - 1, 0) * y ** l + l * x ** k * y ** max(l - 1, 0)
    return s


@pytest.mark.parametrize("diff,l2,linf,tl2,tlinf", [(False, 2.43556e-14, 2.04281e-13, 1e-12, 1e-12),
                                                    (True, 7.02226e-14, 1.13687e-12, 1e-12, 2e-12)])
def test_locproject_tri_golden(diff, l2, linf, tl2, tlinf):
    # -s triangle -b Modified_A Modified_B -o 6 6 -p 7 7 -c 0 0 1 1 0.5 1 [-d]
    b = O.basis_tri(6, 7, 7)
    v = np.array([[[0.0, 0.0], [1.0, 1.0], [0.5, 1.0]]])
    jac, df = O.geom_tri(v, b)
    x, y = O.coords_tri(v, b)
    sol = _tri_sol(x, y, False)
    if diff:
        d0, d1 = O.physderiv_tri(b, 1, False, df, sol)
        sol = d0 + d1
    M = b["M"]
    mass = np.zeros((M, M))
    for i in range(M):
        e = np.zeros(M); e[i] = 1.0
        mass[:, i] = O.iproduct_tri(b, 1, False, jac, O.bwdtrans_tri(b, 1, e))
    coef = np.linalg.solve(mass, O.iproduct_tri(b, 1, False, jac, sol))
    phys = O.bwdtrans_tri(b, 1, coef)
    exact = _tri_sol(x, y, True) if diff else sol
    Wq = np.outer(b["w1"], b["w0"]).ravel()
    got_l2 = np.sqrt(np.sum(Wq * jac[0] * (phys - exact) ** 2))
    got_linf = np.abs(phys - exact).max()
    assert abs(got_l2 - l2) <= tl2
    assert abs(got_linf - linf) <= tlinf
