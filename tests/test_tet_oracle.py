"""Pins of the tetrahedron oracle (collapsed coordinates): Gauss-Radau alpha = 2 points and
Modified_C against independent closed forms, the four operators against dense matrices
from independent numpy code (closed-form bases, gradients from an exact monomial fit),
invariants, and the reference's LocProject tetrahedron goldens
(library/Demos/LocalRegions/Tests/LocProject3D_Tet_Mod_Basis_P6_Q7.tst,
LocProject_Diff3D_Reg_Tet_Mod_Basis_P6_Q7.tst)."""
import numpy as np
import pytest

import dense_ref as R
import oracle as O

REG = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _tets(seed, E=2):
    rng = np.random.default_rng(seed)
    return np.stack([REG + rng.uniform(-0.15, 0.15, (4, 3)) for _ in range(E)])


@pytest.mark.parametrize("nm,nq", [(2, 3), (4, 5), (6, 7), (3, 8)])
def test_modified_c_and_gauss_radau_alpha2(nm, nq):
    b = O.basis_tet(nm, nq + 1, nq, nq)
    z, w = R.gauss_radau(nq, 2)
    assert np.abs(b["z2"] - z).max() < 1e-14
    assert np.abs(b["w2_raw"] - w).max() < 1e-13 * np.abs(w).max()
    C = R.modified_c(nm, z)
    assert np.abs(b["bdata2"].reshape(-1, nq).T - C).max() < 1e-13
    # the derivative table is the exact derivative of the interpolant through the points
    dC = b["dbdata2"].reshape(-1, nq).T
    Dref = np.array([[np.prod([(z[i] - z[c]) for c in range(nq) if c not in (a, i)]) /
                      np.prod([(z[a] - z[c]) for c in range(nq) if c != a]) if a != i else
                      sum(1.0 / (z[i] - z[c]) for c in range(nq) if c != i)
                      for a in range(nq)] for i in range(nq)])
    assert np.abs(dC - Dref @ C).max() < 1e-10 * max(1.0, np.abs(dC).max())


@pytest.mark.parametrize("nm,nq", [(4, 5), (3, 4), (2, 3), (5, 7), (6, 7)])
@pytest.mark.parametrize("deformed", [True, False])
def test_tet_operators_vs_dense(nm, nq, deformed):
    b = O.basis_tet(nm, nq)
    verts = _tets(nm * 10 + nq)
    E = len(verts)
    jac, df = O.geom_tet(verts, b, deformed)
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(nm + nq)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    bw = O.bwdtrans_tet(b, E, u)
    ip = O.iproduct_tet(b, E, deformed, jac, f)
    hh = O.helmholtz_tet(b, E, deformed, jac, df, 2.5, u)
    g = O.physderiv_tet(b, E, deformed, df, bw)
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    for e in range(E):
        m = R.tet_element_mats(nm, nq, nq - 1, nq - 1, verts[e])
        assert m["fit_residual"] < 1e-11  # every mode (with its corrections) is a polynomial
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        assert rel(bw[e * Q:(e + 1) * Q], m["B"] @ ue) < 1e-13
        assert rel(ip[e * M:(e + 1) * M], m["B"].T @ (m["W"] * m["J"] * fe)) < 1e-12
        assert rel(hh[e * M:(e + 1) * M], R.tet_helmholtz_matrix(m, 2.5) @ ue) < 1e-10
        for c in range(3):
            assert rel(g[c][e * Q:(e + 1) * Q], m["Gphys"][c] @ ue) < 1e-10


def test_tet_invariants():
    nm = 5
    b = O.basis_tet(nm, 6)
    verts = _tets(3)
    E = len(verts)
    jac, df = O.geom_tet(verts, b, True)
    M = b["M"]
    # the constant field: modes (0,0,0), (0,0,1), (0,1,0) and (1,0,0)
    one = np.zeros(M)
    one[[0, 1, nm, nm * (nm + 1) // 2]] = 1.0
    ones = np.tile(one, E)
    assert np.abs(O.bwdtrans_tet(b, E, ones) - 1.0).max() < 1e-13
    assert np.abs(O.helmholtz_tet(b, E, True, jac, df, 0.0, ones)).max() < 1e-11
    # the mass of the constant is the volume: |det| * 4/3 per element
    vol = O.iproduct_tet(b, E, True, jac, np.ones(E * b["Q"])) @ ones
    ref = sum(abs(np.linalg.det((v[1:] - v[0]).T)) / 6 for v in verts)
    assert abs(vol - ref) < 1e-13 * ref


def _tet_sol(x, y, z, diff):
    s = np.zeros_like(x)
    for k in range(6):
        for l in range(6 - k):
            for m in range(6 - k - l):
                if not diff:
                    s += x ** k * y ** l * z ** m
                else:
                    s += (k * x ** max(k - 1, 0) * y ** l * z ** m + l * x ** k * y ** max(l - 1, 0) * z ** m
                          + m * x ** k * y ** l * z ** max(m - 1, 0))
    return s


@pytest.mark.parametrize("diff,l2,linf,tl2,tlinf", [(False, 1.89448e-13, 1.17817e-12, 1e-12, 1e-11),
                                                    (True, 9.02964e-13, 5.08749e-12, 1e-12, 1e-11)])
def test_locproject_tet_golden(diff, l2, linf, tl2, tlinf):
    # -s tetrahedron -b Modified_A Modified_B Modified_C -o 6 6 6 -p 7 7 7 -c <REG> [-d]
    b = O.basis_tet(6, 7, 7, 7)
    v = REG[None]
    jac, df = O.geom_tet(v, b, False)
    x, y, z = O.coords_tet(v, b)
    sol = _tet_sol(x, y, z, False)
    if diff:
        sol = sum(O.physderiv_tet(b, 1, False, df, sol))
    M = b["M"]
    mass = np.zeros((M, M))
    for i in range(M):
        e = np.zeros(M); e[i] = 1.0
        mass[:, i] = O.iproduct_tet(b, 1, False, jac, O.bwdtrans_tet(b, 1, e))
    coef = np.linalg.solve(mass, O.iproduct_tet(b, 1, False, jac, sol))
    phys = O.bwdtrans_tet(b, 1, coef)
    exact = _tet_sol(x, y, z, True) if diff else sol
    Wq = np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel()
    got_l2 = np.sqrt(np.sum(Wq * jac[0] * (phys - exact) ** 2))
    got_linf = np.abs(phys - exact).max()
    print(f"tet locproject diff={diff}: L2 {got_l2:.3e} Linf {got_linf:.3e}")
    assert abs(got_l2 - l2) <= tl2
    assert abs(got_linf - linf) <= tlinf
