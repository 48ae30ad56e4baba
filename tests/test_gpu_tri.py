"""Triangle operators on the GPU (collapsed coordinates, FP64) against the CPU oracle,
which is pinned to dense collapsed-coordinate matrices and to the reference's LocProject
triangle goldens (tests/test_tri_oracle.py).  Bar: 1e-12 relative."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O

pytestmark = pytest.mark.gpu

COMBOS = [(6, 7), (3, 5), (2, 3), (4, 8), (8, 16), (5, 6)]


def _mesh(E, seed):
    """E straight-sided triangles of random shape and orientation (positive area)."""
    rng = np.random.default_rng(seed)
    v = np.zeros((E, 3, 2))
    for e in range(E):
        c = rng.uniform(-2, 2, 2)
        while True:
            t = c + rng.uniform(-0.6, 0.6, (3, 2))
            d1, d2 = t[1] - t[0], t[2] - t[0]
            if d1[0] * d2[1] - d1[1] * d2[0] > 0.05:
                break
        v[e] = t
    return v


def _tables(b):
    a = dict(bdata=b["bdata0"], dbdata=b["dbdata0"], D=b["D0"], w=b["w0"])
    c = dict(bdata=b["bdata1"], dbdata=b["dbdata1"], D=b["D1"], w=b["w1_raw"])
    return [a, c]


def _rel(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


@pytest.mark.parametrize("nm,nq0", COMBOS)
@pytest.mark.parametrize("deformed", [False, True])
def test_tri_operators_match_the_oracle(nm, nq0, deformed):
    print(f"[case] tri nm={nm} nq0={nq0} deformed={deformed}", flush=True)
    E = 301
    b = O.basis_tri(nm, nq0)
    verts = _mesh(E, nm * 100 + nq0)
    jac, df = O.geom_tri(verts, b, deformed)
    # the GPU geometry kernel gives the same factors
    gj, gd = H.geom_device("tri", torch.from_numpy(verts).cuda(), (nq0, nq0 - 1), None, deformed)
    assert np.abs(gj.cpu().numpy() - jac).max() < 1e-13 * np.abs(jac).max()
    assert np.abs(gd.cpu().numpy() - df).max() < 1e-13 * np.abs(df).max()
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    geo = "Deformed" if deformed else "Regular"
    tabs = _tables(b)
    du, dfq = torch.from_numpy(u).cuda(), torch.from_numpy(f).cuda()

    bw = H.Operator("BwdTrans_Tri_Regular", nm, nq0, E, basis=tabs)
    got = bw.apply(du).cpu().numpy()
    assert _rel(got, O.bwdtrans_tri(b, E, u)) < 1e-12

    ip = H.Operator(f"IProduct_Tri_{geo}", nm, nq0, E, basis=tabs)
    ip.set_jac(jac)
    got = ip.apply(dfq).cpu().numpy()
    assert _rel(got, O.iproduct_tri(b, E, deformed, jac, f)) < 1e-12

    hm = H.Operator(f"Helmholtz_Tri_{geo}", nm, nq0, E, basis=tabs)
    hm.set_jac(jac)
    hm.set_df(df)
    hm.set_lambda(2.5)
    got = hm.apply(du).cpu().numpy()
    assert _rel(got, O.helmholtz_tri(b, E, deformed, jac, df, 2.5, u)) < 1e-12

    pd = H.Operator(f"PhysDeriv_Tri_{geo}", nm, nq0, E, basis=tabs)
    pd.set_df(df)
    g = pd.apply(dfq).cpu().numpy()
    r0, r1 = O.physderiv_tri(b, E, deformed, df, f)
    assert _rel(g[:E * Q], r0) < 1e-12 and _rel(g[E * Q:], r1) < 1e-12

    ipd = H.Operator(f"IProductWRTDerivBase_Tri_{geo}", nm, nq0, E, basis=tabs)
    ipd.set_jac(jac)
    ipd.set_df(df)
    fs = [rng.uniform(-1, 1, E * Q) for _ in range(2)]
    dfs = torch.from_numpy(np.concatenate(fs)).cuda()
    got = ipd.apply(dfs).cpu().numpy()
    assert _rel(got, O.iprodderiv_tri(b, E, deformed, jac, df, fs[0], fs[1])) < 1e-12
    with pytest.raises(H.MFError):
        ipd.apply_transpose(du)

    # Euclidean transposes: <A x, y> = <x, A^T y>
    for op, nin, nout in [(bw, M, Q), (ip, Q, M), (hm, M, M)]:
        x = torch.from_numpy(rng.uniform(-1, 1, E * nin)).cuda()
        y = torch.from_numpy(rng.uniform(-1, 1, E * nout)).cuda()
        lhs = torch.dot(op.apply(x), y).item()
        rhs = torch.dot(x, op.apply_transpose(y)).item()
        assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_tri_locproject_golden_on_the_gpu():
    """LocProject2D_Tri_Mod_Basis_P6_Q7.tst (L2 2.43556e-14, Linf 2.04281e-13, tolerance
    1e-12) with the GPU BwdTrans / IProduct forming the projection."""
    print("[case] tri locproject", flush=True)
    b = O.basis_tri(6, 7, 7)
    tabs = _tables(b)
    v = np.array([[[0.0, 0.0], [1.0, 1.0], [0.5, 1.0]]])
    jac, _ = O.geom_tri(v, b)
    x, y = O.coords_tri(v, b)
    sol = sum(x ** k * y ** l for k in range(6) for l in range(6 - k))
    bw = H.Operator("BwdTrans_Tri_Regular", 6, (7, 7), 1, basis=tabs)
    ip = H.Operator("IProduct_Tri_Regular", 6, (7, 7), 1, basis=tabs)
    ip.set_jac(jac)
    M = b["M"]
    eye = torch.eye(M, dtype=torch.float64).cuda()
    mass = torch.stack([ip.apply(bw.apply(eye[i].contiguous())) for i in range(M)], 1)
    rhs = ip.apply(torch.from_numpy(sol).cuda())
    coef = np.linalg.solve(mass.cpu().numpy(), rhs.cpu().numpy())
    phys = bw.apply(torch.from_numpy(coef).cuda()).cpu().numpy()
    Wq = np.outer(b["w1"], b["w0"]).ravel()
    l2 = np.sqrt(np.sum(Wq * jac[0] * (phys - sol) ** 2))
    assert abs(l2 - 2.43556e-14) <= 1e-12
    assert abs(np.abs(phys - sol).max() - 2.04281e-13) <= 1e-12


def test_tri_product_tables_equal_the_oracle_tables_on_the_gpu_path():
    b = O.basis_tri(6, 7)
    E = 64
    verts = _mesh(E, 3)
    jac, df = O.geom_tri(verts, b)
    u = np.random.default_rng(1).uniform(-1, 1, E * b["M"])
    hm = H.Operator("Helmholtz_Tri_Regular", 6, 7, E)  # library-generated tables
    hm.set_jac(jac)
    hm.set_df(df)
    hm.set_lambda(1.0)
    got = hm.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert _rel(got, O.helmholtz_tri(b, E, False, jac, df, 1.0, u)) < 1e-12


def test_tri_rejects_what_it_does_not_cover():
    b = O.basis_tri(6, 7)
    tabs = _tables(b)
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Tri_Regular", 6, 7, 4, precision=H.FP32, basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("IProductWRTDerivBase_Tri_Regular", 6, 7, 4, precision=H.FP32,
                   basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Tri_Regular", 9, 10, 4)
    op = H.Operator("PhysDeriv_Tri_Regular", 6, 7, 4, basis=tabs)
    op.set_df(np.zeros(16))
    with pytest.raises(H.MFError):
        op.apply_transpose(torch.zeros(2 * 4 * 42, dtype=torch.float64).cuda())
