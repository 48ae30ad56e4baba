"""Pins of the pyramid oracle (collapsed coordinates): ModifiedPyr_C against closed forms,
the four operators against dense matrices from independent numpy code (closed-form modes
as functions of xi, exact complex-step derivatives of the modes and of the map, no
collapsed chain rule shared with the oracle), invariants, and the reference's LocProject
pyramid goldens (library/Demos/LocalRegions/Tests/LocProject3D_Pyr_Mod_Basis_P6_Q7.tst,
LocProject_Diff3D_Reg_Pyr_Mod_Basis_P6_Q7.tst)."""
import numpy as np
import pytest
from scipy.special import eval_jacobi

import dense_ref as R
import oracle as O

REG = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
                [0.5, 0.5, 0.866]])


def _pyrs(seed, E=2):
    rng = np.random.default_rng(seed)
    return np.stack([REG + rng.uniform(-0.12, 0.12, (5, 3)) for _ in range(E)])


def test_complex_jacobi_matches_scipy():
    x = np.linspace(-1, 1, 17)
    for n in range(7):
        for a, b in [(1, 1), (2, 1), (5, 1), (0, 0)]:
            assert np.allclose(R.jacobi_c(n, a, b, x), eval_jacobi(n, a, b, x), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("nm,nq", [(2, 2), (4, 5), (6, 6), (3, 8)])
def test_modified_pyr_c_closed_form(nm, nq):
    b = O.basis_pyr(nm, nq + 1, nq)
    z, w = R.gauss_radau(nq, 2)
    assert np.abs(b["z2"] - z).max() < 1e-14
    C = np.stack([R.pyr_c(p, q, r, z) for p, q, r in R.pyr_modes(nm)], 1)
    assert np.abs(b["bdata2"].reshape(-1, nq).T - C).max() < 1e-13


@pytest.mark.parametrize("nm,nq", [(4, 5), (3, 4), (2, 3), (5, 7), (6, 7)])
@pytest.mark.parametrize("deformed", [True, False])
def test_pyr_operators_vs_dense(nm, nq, deformed):
    b = O.basis_pyr(nm, nq)
    # regular: a pyramid on a parallelogram base (affine map); deformed: perturbed vertices
    verts = _pyrs(nm * 10 + nq) if deformed else np.stack([REG, REG * 1.3 + 0.2])
    E = len(verts)
    jac, df = O.geom_pyr(verts, b, deformed)
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(nm + nq)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    bw = O.bwdtrans_pyr(b, E, u)
    ip = O.iproduct_pyr(b, E, deformed, jac, f)
    hh = O.helmholtz_pyr(b, E, deformed, jac, df, 2.5, u)
    g = O.physderiv_pyr(b, E, deformed, df, bw)
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    for e in range(E):
        m = R.pyr_element_mats(nm, nq, nq, nq - 1, verts[e])
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        if deformed:
            assert np.abs(jac[e * Q:(e + 1) * Q] - m["J"]).max() < 1e-13 * m["J"].max()
        assert rel(bw[e * Q:(e + 1) * Q], m["B"] @ ue) < 1e-13
        assert rel(ip[e * M:(e + 1) * M], m["B"].T @ (m["W"] * m["J"] * fe)) < 1e-12
        assert rel(hh[e * M:(e + 1) * M], R.pyr_helmholtz_matrix(m, 2.5) @ ue) < 1e-11
        for c in range(3):
            assert rel(g[c][e * Q:(e + 1) * Q], m["Gphys"][c] @ ue) < 1e-11


def test_collocation_gradient_is_exact_when_the_points_resolve_every_mode():
    """With nq2 >= 2 nm - 3 points along eta2 every mode is resolved, and the collocation
    gradient (the reference's) equals the exact one; with fewer points it does not."""
    nm, v = 4, _pyrs(11)[0]
    exact = R.pyr_element_mats(nm, nm + 1, nm + 1, 2 * nm - 3, v, deriv="exact")
    colloc = R.pyr_element_mats(nm, nm + 1, nm + 1, 2 * nm - 3, v)
    for c in range(3):
        assert np.abs(exact["Gphys"][c] - colloc["Gphys"][c]).max() < 1e-9
    exact = R.pyr_element_mats(nm, nm + 1, nm + 1, nm, v, deriv="exact")
    colloc = R.pyr_element_mats(nm, nm + 1, nm + 1, nm, v)
    assert max(np.abs(exact["Gphys"][c] - colloc["Gphys"][c]).max() for c in range(3)) > 1e-6


def test_pyr_invariants():
    nm = 5
    b = O.basis_pyr(nm, 6)
    verts = _pyrs(3)
    E = len(verts)
    jac, df = O.geom_pyr(verts, b, True)
    M = b["M"]
    T = nm * (nm + 1) // 2
    # the constant field: the five vertex modes (0,0,0), (0,0,1), (0,1,0), (1,0,0), (1,1,0)
    one = np.zeros(M)
    one[[0, 1, nm, T, T + nm - 1]] = 1.0
    ones = np.tile(one, E)
    assert np.abs(O.bwdtrans_pyr(b, E, ones) - 1.0).max() < 1e-13
    assert np.abs(O.helmholtz_pyr(b, E, True, jac, df, 0.0, ones)).max() < 1e-11
    # x, y, z themselves are in the space: their gradients are the unit vectors
    x = O.coords_pyr(verts, b)
    for c in range(3):
        g = O.physderiv_pyr(b, E, True, df, x[c])
        for a in range(3):
            assert np.abs(g[a] - (1.0 if a == c else 0.0)).max() < 1e-11


def _sol(x, y, z, diff):
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


@pytest.mark.parametrize("diff,l2,linf,tol", [(False, 1.15445e-12, 8.91376e-12, 1e-8),
                                              (True, 4.64852e-12, 2.413e-11, 1e-9)])
def test_locproject_pyr_golden(diff, l2, linf, tol):
    # -s pyramid -b Modified_A Modified_A ModifiedPyr_C -o 6 6 6 -p 7 7 6 -c <REG> [-d]
    b = O.basis_pyr(6, 7, 6)
    v = REG[None]
    jac, df = O.geom_pyr(v, b, False)
    x, y, z = O.coords_pyr(v, b)
    sol = _sol(x, y, z, False)
    if diff:
        sol = sum(O.physderiv_pyr(b, 1, False, df, sol))
    M = b["M"]
    mass = np.zeros((M, M))
    for i in range(M):
        e = np.zeros(M); e[i] = 1.0
        mass[:, i] = O.iproduct_pyr(b, 1, False, jac, O.bwdtrans_pyr(b, 1, e))
    coef = np.linalg.solve(mass, O.iproduct_pyr(b, 1, False, jac, sol))
    phys = O.bwdtrans_pyr(b, 1, coef)
    exact = _sol(x, y, z, True) if diff else sol
    Wq = np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel()
    got_l2 = np.sqrt(np.sum(Wq * jac[0] * (phys - exact) ** 2))
    got_linf = np.abs(phys - exact).max()
    print(f"pyr locproject diff={diff}: L2 {got_l2:.3e} Linf {got_linf:.3e}")
    assert abs(got_l2 - l2) <= tol
    assert abs(got_linf - linf) <= tol
