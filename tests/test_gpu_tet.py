"""Tetrahedron operators on the GPU (collapsed coordinates, FP64) against the CPU oracle,
which is pinned to dense collapsed-coordinate matrices and to the reference's LocProject
tetrahedron goldens (tests/test_tet_oracle.py).  Bar: 1e-12 relative."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O

pytestmark = pytest.mark.gpu

REG = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
# (nm, (nq0, nq1, nq2)): the reference's nq1 = nq2 = nq0 - 1 plus other point counts
COMBOS = [(4, (5, 4, 4)), (2, (3, 2, 2)), (3, (4, 3, 3)), (6, (7, 6, 6)), (8, (9, 8, 8)),
          (5, (7, 6, 5))]


def _mesh(E, seed):
    """E tetrahedra of random shape, size and position, positively oriented."""
    rng = np.random.default_rng(seed)
    v = np.zeros((E, 4, 3))
    for e in range(E):
        while True:
            t = REG * rng.uniform(0.5, 2.0) + rng.uniform(-0.25, 0.25, (4, 3)) + rng.uniform(-3, 3, 3)
            if np.linalg.det((t[1:] - t[0]).T) > 0.05:
                break
        v[e] = t
    return v


def _basis(nm, nq):
    return O.basis_tet(nm, nq[0], nq[1], nq[2])


def _tables(b):
    return [dict(bdata=b["bdata0"], dbdata=b["dbdata0"], D=b["D0"], w=b["w0"]),
            dict(bdata=b["bdata1"], dbdata=b["dbdata1"], D=b["D1"], w=b["w1_raw"]),
            dict(bdata=b["bdata2"], dbdata=b["dbdata2"], D=b["D2"], w=b["w2_raw"])]


def _rel(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


@pytest.mark.parametrize("nm,nq", COMBOS)
@pytest.mark.parametrize("deformed", [False, True])
def test_tet_operators_match_the_oracle(nm, nq, deformed):
    print(f"[case] tet nm={nm} nq={nq} deformed={deformed}", flush=True)
    E = 301
    b = _basis(nm, nq)
    verts = _mesh(E, nm * 100 + nq[0])
    jac, df = O.geom_tet(verts, b, deformed)
    gj, gd = H.geom_device("tet", torch.from_numpy(verts).cuda(), nq, None, deformed)
    assert np.abs(gj.cpu().numpy() - jac).max() < 1e-13 * np.abs(jac).max()
    assert np.abs(gd.cpu().numpy() - df).max() < 1e-13 * np.abs(df).max()
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    geo = "Deformed" if deformed else "Regular"
    tabs = _tables(b)
    du, dfq = torch.from_numpy(u).cuda(), torch.from_numpy(f).cuda()

    bw = H.Operator("BwdTrans_Tet_Regular", nm, nq, E, basis=tabs)
    assert _rel(bw.apply(du).cpu().numpy(), O.bwdtrans_tet(b, E, u)) < 1e-12

    ip = H.Operator(f"IProduct_Tet_{geo}", nm, nq, E, basis=tabs)
    ip.set_jac(gj)
    assert _rel(ip.apply(dfq).cpu().numpy(), O.iproduct_tet(b, E, deformed, jac, f)) < 1e-12

    hm = H.Operator(f"Helmholtz_Tet_{geo}", nm, nq, E, basis=tabs)
    hm.set_jac(gj)
    hm.set_df(gd)
    hm.set_lambda(2.5)
    got = hm.apply(du).cpu().numpy()
    assert _rel(got, O.helmholtz_tet(b, E, deformed, jac, df, 2.5, u)) < 1e-12

    pd = H.Operator(f"PhysDeriv_Tet_{geo}", nm, nq, E, basis=tabs)
    pd.set_df(gd)
    g = pd.apply(dfq).cpu().numpy()
    ref = O.physderiv_tet(b, E, deformed, df, f)
    for c in range(3):
        assert _rel(g[c * E * Q:(c + 1) * E * Q], ref[c]) < 1e-12

    ipd = H.Operator(f"IProductWRTDerivBase_Tet_{geo}", nm, nq, E, basis=tabs)
    ipd.set_jac(gj)
    ipd.set_df(gd)
    fs = [rng.uniform(-1, 1, E * Q) for _ in range(3)]
    dfs = torch.from_numpy(np.concatenate(fs)).cuda()
    got = ipd.apply(dfs).cpu().numpy()
    assert _rel(got, O.iprodderiv_tet(b, E, deformed, jac, df, fs)) < 1e-12
    with pytest.raises(H.MFError):
        ipd.apply_transpose(du)

    # Euclidean transposes: <A x, y> = <x, A^T y>
    for op, nin, nout in [(bw, M, Q), (ip, Q, M), (hm, M, M)]:
        x = torch.from_numpy(rng.uniform(-1, 1, E * nin)).cuda()
        y = torch.from_numpy(rng.uniform(-1, 1, E * nout)).cuda()
        lhs = torch.dot(op.apply(x), y).item()
        rhs = torch.dot(x, op.apply_transpose(y)).item()
        assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_tet_many_elements_per_workgroup_loop():
    """More elements than workgroups, with library-generated tables."""
    print("[case] tet E=20000", flush=True)
    nm, nq, E = 3, (4, 3, 3), 20000
    b = _basis(nm, nq)
    verts = _mesh(E, 5)
    jac, df = O.geom_tet(verts, b, True)
    u = np.random.default_rng(2).uniform(-1, 1, E * b["M"])
    hm = H.Operator("Helmholtz_Tet_Deformed", nm, nq, E)
    hm.set_jac(jac)
    hm.set_df(df)
    hm.set_lambda(0.5)
    got = hm.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert _rel(got, O.helmholtz_tet(b, E, True, jac, df, 0.5, u)) < 1e-12


def _sol(x, y, z, diff):
    s = np.zeros_like(x)
    for k in range(6):
        for l in range(6 - k):
            for m in range(6 - k - l):
                if not diff:
                    s += x ** k * y ** l * z ** m
                else:
                    s += (k * x ** max(k - 1, 0) * y ** l * z ** m
                          + l * x ** k * y ** max(l - 1, 0) * z ** m
                          + m * x ** k * y ** l * z ** max(m - 1, 0))
    return s


@pytest.mark.parametrize("diff,l2,linf", [(False, 1.89448e-13, 1.17817e-12),
                                          (True, 9.02964e-13, 5.08749e-12)])
def test_tet_locproject_golden_on_the_gpu(diff, l2, linf):
    """LocProject3D_Tet_Mod_Basis_P6_Q7.tst and LocProject_Diff3D_Reg_Tet_Mod_Basis_P6_Q7.tst
    (tolerances 1e-12 on L2, 1e-11 on Linf) with the GPU operators forming the projection."""
    print(f"[case] tet locproject diff={diff}", flush=True)
    nq = (7, 7, 7)
    b = _basis(6, nq)
    tabs = _tables(b)
    v = REG[None]
    gj, gd = H.geom_device("tet", torch.from_numpy(v).cuda(), nq, None, False)
    x, y, z = O.coords_tet(v, b)
    sol = _sol(x, y, z, False)
    if diff:
        pd = H.Operator("PhysDeriv_Tet_Regular", 6, nq, 1, basis=tabs)
        pd.set_df(gd)
        sol = pd.apply(torch.from_numpy(sol).cuda()).cpu().numpy().reshape(3, -1).sum(0)
    bw = H.Operator("BwdTrans_Tet_Regular", 6, nq, 1, basis=tabs)
    ip = H.Operator("IProduct_Tet_Regular", 6, nq, 1, basis=tabs)
    ip.set_jac(gj)
    M = b["M"]
    eye = torch.eye(M, dtype=torch.float64).cuda()
    mass = torch.stack([ip.apply(bw.apply(eye[i].contiguous())) for i in range(M)], 1)
    coef = np.linalg.solve(mass.cpu().numpy(), ip.apply(torch.from_numpy(sol).cuda()).cpu().numpy())
    phys = bw.apply(torch.from_numpy(coef).cuda()).cpu().numpy()
    exact = _sol(x, y, z, True) if diff else sol
    Wq = np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel()
    jac = gj.cpu().numpy()[0]
    assert abs(np.sqrt(np.sum(Wq * jac * (phys - exact) ** 2)) - l2) <= 1e-12
    assert abs(np.abs(phys - exact).max() - linf) <= 1e-11


def test_tet_rejects_what_it_does_not_cover():
    tabs = _tables(O.basis_tet(4, 5))
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Tet_Regular", 4, 5, 4, precision=H.FP32, basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("IProductWRTDerivBase_Tet_Regular", 4, 5, 4, precision=H.FP32,
                   basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Tet_Regular", 9, 10, 4)
    with pytest.raises(H.MFError):  # one element's arrays exceed a workgroup's LDS
        H.Operator("Helmholtz_Tet_Regular", 8, (16, 16, 16), 4)
    op = H.Operator("PhysDeriv_Tet_Regular", 4, 5, 4, basis=tabs)
    op.set_df(np.zeros(36))
    with pytest.raises(H.MFError):
        op.apply_transpose(torch.zeros(3 * 4 * 80, dtype=torch.float64).cuda())
