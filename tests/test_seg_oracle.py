"""The oracle's segment operators (oracle/mf_oracle_seg.c) pinned by closed forms of the
Modified_A basis on GLL points: phi_0 = (1 - xi)/2, phi_1 = (1 + xi)/2, phi_p = phi_0 phi_1
P_{p-2}^{(1,1)}(xi), so that the integrals, boundary values and derivatives below are known
exactly.  Each planted mistake (a missing Jacobian, a transposed table, a sign, a dropped
coordinate) fails one of them."""
import numpy as np
import pytest

import oracle as O

SEGS = {1: np.array([[[-0.3], [0.9]], [[2.0], [1.2]]]),       # the second one reversed
        2: np.array([[[-1.5, 0.0], [1.0, 1.0]], [[0.2, 0.4], [-0.6, 1.1]]]),
        3: np.array([[[0.0, 0.0, 0.0], [1.0, 2.0, 2.0]], [[1.0, -1.0, 0.5], [0.2, 0.3, -0.4]]])}


def _t(verts):
    return 0.5 * (verts[:, 1] - verts[:, 0])


@pytest.mark.parametrize("ncoord", [1, 2, 3])
@pytest.mark.parametrize("deformed", [False, True])
def test_seg_geometry_is_the_half_chord(ncoord, deformed):
    v = SEGS[ncoord]
    nq = 5
    jac, df = O.geom_seg(v, nq, deformed)
    t = _t(v)
    L = np.linalg.norm(t, axis=1)
    rep = nq if deformed else 1
    assert np.allclose(jac, np.repeat(L, rep), rtol=1e-15)
    df = df.reshape(ncoord, -1)
    for c in range(ncoord):
        assert np.allclose(df[c], np.repeat(t[:, c] / L ** 2, rep), rtol=1e-15)
    if ncoord == 2:  # the segment (-1.5, 0) -> (1, 1): t = (1.25, 0.5), J^2 = 1.8125
        assert abs(jac[0] ** 2 - 1.8125) < 1e-14


@pytest.mark.parametrize("nm,nq", [(2, 2), (4, 5), (7, 9), (16, 20)])
def test_seg_bwdtrans_gives_the_vertex_and_bubble_modes(nm, nq):
    b = O.basis(nm, nq)
    z = b["z"]
    E = 3
    u = np.zeros(E * nm)
    u[0] = u[1] = 1.0           # element 0: phi_0 + phi_1 = 1
    u[nm + 1] = 1.0             # element 1: phi_1
    if nm > 2:
        u[2 * nm + 2] = 1.0     # element 2: phi_2 = (1 - xi^2)/4
    out = O.bwdtrans_seg(b, E, u).reshape(E, nq)
    assert np.allclose(out[0], 1.0, atol=1e-14)
    assert np.allclose(out[1], 0.5 * (1 + z), atol=1e-14)
    if nm > 2:
        assert np.allclose(out[2], 0.25 * (1 - z * z), atol=1e-14)


@pytest.mark.parametrize("nm,nq", [(4, 5), (7, 9), (16, 20)])
@pytest.mark.parametrize("deformed", [False, True])
def test_seg_iproduct_of_one_is_length_over_the_modes(nm, nq, deformed):
    b = O.basis(nm, nq)
    v = SEGS[2]
    E = v.shape[0]
    jac, _ = O.geom_seg(v, nq, deformed)
    L = np.linalg.norm(_t(v), axis=1)
    out = O.iproduct_seg(b, E, deformed, jac, np.ones(E * nq)).reshape(E, nm)
    # int phi_0 = int phi_1 = J, int phi_2 = J/3, higher modes orthogonal to 1 in (1-xi^2)
    ref = np.zeros((E, nm))
    ref[:, 0] = ref[:, 1] = L
    ref[:, 2] = L / 3
    assert np.allclose(out, ref, atol=1e-14)


@pytest.mark.parametrize("ncoord", [1, 2, 3])
@pytest.mark.parametrize("deformed", [False, True])
def test_seg_physderiv_of_a_cubic(ncoord, deformed):
    nq = 6
    b = O.basis(4, nq)
    z = b["z"]
    v = SEGS[ncoord]
    E = v.shape[0]
    _, df = O.geom_seg(v, nq, deformed)
    t = _t(v)
    g = (t ** 2).sum(1)
    u = np.tile(z ** 3 - 2 * z, E)
    out = O.physderiv_seg(b, E, deformed, ncoord, df, u).reshape(ncoord, E, nq)
    for c in range(ncoord):
        for e in range(E):
            assert np.allclose(out[c, e], (3 * z ** 2 - 2) * t[e, c] / g[e], atol=1e-12)


@pytest.mark.parametrize("ncoord", [1, 2, 3])
@pytest.mark.parametrize("deformed", [False, True])
def test_seg_iprodderiv_of_the_unit_tangent_is_the_boundary_values(ncoord, deformed):
    """in_c = t_c/|t| (unit tangent): sum_c df_c in_c = 1/J, so the operator gives
    sum_i w_i dphi_p/dxi(z_i) = phi_p(1) - phi_p(-1): -1, +1 and zeros."""
    nm, nq = 6, 8
    b = O.basis(nm, nq)
    v = SEGS[ncoord]
    E = v.shape[0]
    jac, df = O.geom_seg(v, nq, deformed)
    t = _t(v)
    tu = t / np.linalg.norm(t, axis=1)[:, None]
    f = np.concatenate([np.repeat(tu[:, c], nq) for c in range(ncoord)])
    out = O.iprodderiv_seg(b, E, deformed, ncoord, jac, df, f).reshape(E, nm)
    ref = np.zeros((E, nm))
    ref[:, 0], ref[:, 1] = -1.0, 1.0
    assert np.allclose(out, ref, atol=1e-13)


@pytest.mark.parametrize("ncoord", [1, 2, 3])
def test_seg_iprodderiv_is_the_adjoint_of_the_derivative(ncoord):
    """<IPDB(f), c> = sum_c sum_i w_i J_i f_c,i (PhysDeriv(BwdTrans(c)))_c,i: the left side
    uses the derivative table dB, the right side the collocation matrix D."""
    nm, nq = 7, 9
    b = O.basis(nm, nq)
    rng = np.random.default_rng(ncoord)
    E = 5
    v = rng.uniform(-1, 1, (E, 2, ncoord))
    jac, df = O.geom_seg(v, nq, True)
    f = rng.uniform(-1, 1, ncoord * E * nq)
    c = rng.uniform(-1, 1, E * nm)
    lhs = O.iprodderiv_seg(b, E, True, ncoord, jac, df, f) @ c
    du = O.physderiv_seg(b, E, True, ncoord, df, O.bwdtrans_seg(b, E, c))
    wj = np.tile(np.tile(b["w"], E) * jac, ncoord)
    assert abs(lhs - np.sum(wj * f * du)) < 1e-12 * max(1.0, abs(lhs))
