"""Pins of the prism oracle (collapsed coordinates): the four operators against dense
matrices from independent numpy code (closed-form bases, gradients from an exact
monomial fit), invariants, and the reference's LocProject prism goldens
(library/Demos/LocalRegions/Tests/LocProject3D_Prism_Mod_Basis_P6_Q7.tst,
LocProject_Diff3D_Reg_Prism_Mod_Basis_P6_Q7.tst)."""
import numpy as np
import pytest

import dense_ref as R
import oracle as O

REG = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
                [0.5, 0.0, 1.0], [0.5, 1.0, 1.0]])


def _prisms(seed, E=2):
    rng = np.random.default_rng(seed)
    return np.stack([REG + rng.uniform(-0.12, 0.12, (6, 3)) for _ in range(E)])


@pytest.mark.parametrize("nm,nq", [(4, 5), (3, 4), (2, 3), (5, 7)])
@pytest.mark.parametrize("deformed", [True, False])
def test_prism_operators_vs_dense(nm, nq, deformed):
    b = O.basis_prism(nm, nq)
    verts = _prisms(nm * 10 + nq) if deformed else np.stack([REG, REG * 1.3 + 0.2])
    E = len(verts)
    jac, df = O.geom_prism(verts, b, deformed)
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(nm + nq)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    bw = O.bwdtrans_prism(b, E, u)
    ip = O.iproduct_prism(b, E, deformed, jac, f)
    hh = O.helmholtz_prism(b, E, deformed, jac, df, 2.5, u)
    g = O.physderiv_prism(b, E, deformed, df, bw)
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    for e in range(E):
        m = R.prism_element_mats(nm, nq, nq - 1, verts[e])
        assert m["fit_residual"] < 1e-11
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        assert rel(bw[e * Q:(e + 1) * Q], m["B"] @ ue) < 1e-13
        assert rel(ip[e * M:(e + 1) * M], m["B"].T @ (m["W"] * m["J"] * fe)) < 1e-12
        assert rel(hh[e * M:(e + 1) * M], R.prism_helmholtz_matrix(m, 2.5) @ ue) < 1e-10
        for c in range(3):
            assert rel(g[c][e * Q:(e + 1) * Q], m["Gphys"][c] @ ue) < 1e-10


def test_prism_invariants():
    b = O.basis_prism(5, 6)
    verts = _prisms(3)
    E = len(verts)
    jac, df = O.geom_prism(verts, b, True)
    M = b["M"]
    # the constant field is the sum of the four bottom vertex modes (p, q, 0), p, q < 2,
    # and the two top singular-edge modes (0, q, 1), q < 2
    def mode(p, q, r):
        return sum(5 * (5 - pp) for pp in range(p)) + q * (5 - p) + r
    one = np.zeros(M)
    for p in range(2):
        for q in range(2):
            one[mode(p, q, 0)] = 1.0
    for q in range(2):
        one[mode(0, q, 1)] = 1.0
    ones = np.tile(one, E)
    assert np.abs(O.bwdtrans_prism(b, E, ones) - 1.0).max() < 1e-13
    assert np.abs(O.helmholtz_prism(b, E, True, jac, df, 0.0, ones)).max() < 1e-11
    Hu = O.helmholtz_prism(b, E, True, jac, df, 3.0, np.random.default_rng(0).uniform(-1, 1, E * M))
    assert np.all(np.isfinite(Hu))


def _prism_sol(x, y, z, diff):
    s = np.zeros_like(x)
    for k in range(6):
        for l in range(6):
            for m in range(6 - k):
                if not diff:
                    s += x ** k * y ** l * z ** m
                else:
                    s += (k * x ** max(k - 1, 0) * y ** l * z ** m + l * x ** k * y ** max(l - 1, 0) * z ** m
                          + m * x ** k * y ** l * z ** max(m - 1, 0))
    return s


@pytest.mark.parametrize("diff,l2,linf,tl2,tlinf", [(False, 5.16422e-12, 2.08524e-11, 1e-8, 1e-8),
                                                    (True, 2.41945e-11, 1.03926e-10, 1e-9, 1e-9)])
def test_locproject_prism_golden(diff, l2, linf, tl2, tlinf):
    # -s prism -b Modified_A Modified_A Modified_B -o 6 6 6 -p 7 7 7 -c <REG> [-d]
    b = O.basis_prism(6, 7, 7)
    v = REG[None]
    jac, df = O.geom_prism(v, b, False)
    x, y, z = O.coords_prism(v, b)
    sol = _prism_sol(x, y, z, False)
    if diff:
        sol = sum(O.physderiv_prism(b, 1, False, df, sol))
    M = b["M"]
    mass = np.zeros((M, M))
    for i in range(M):
        e = np.zeros(M); e[i] = 1.0
        mass[:, i] = O.iproduct_prism(b, 1, False, jac, O.bwdtrans_prism(b, 1, e))
    coef = np.linalg.solve(mass, O.iproduct_prism(b, 1, False, jac, sol))
    phys = O.bwdtrans_prism(b, 1, coef)
    exact = _prism_sol(x, y, z, True) if diff else sol
    Wq = np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel()
    got_l2 = np.sqrt(np.sum(Wq * jac[0] * (phys - exact) ** 2))
    got_linf = np.abs(phys - exact).max()
    print(f"prism locproject diff={diff}: L2 {got_l2:.3e} Linf {got_linf:.3e}")
    assert abs(got_l2 - l2) <= tl2
    assert abs(got_linf - linf) <= tlinf
