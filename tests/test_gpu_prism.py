"""Prism operators on the GPU (collapsed coordinates, FP64) against the CPU oracle, which
is pinned to dense collapsed-coordinate matrices and to the reference's LocProject prism
goldens (tests/test_prism_oracle.py).  Bar: 1e-12 relative."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O

pytestmark = pytest.mark.gpu

REG = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
                [0.5, 0.0, 1.0], [0.5, 1.0, 1.0]])
# (nm, (nq0, nq1, nq2)): the reference's nq2 = nq0 - 1 pairs plus other point counts
COMBOS = [(4, (5, 5, 4)), (2, (3, 3, 2)), (3, (4, 4, 3)), (6, (7, 7, 6)), (8, (9, 9, 8)),
          (5, (7, 6, 6))]


def _mesh(E, seed, deformed):
    """E prisms: perturbed copies of REG (deformed) or scaled and shifted ones (affine)."""
    rng = np.random.default_rng(seed)
    if deformed:
        return np.stack([REG + rng.uniform(-0.12, 0.12, (6, 3)) for _ in range(E)])
    return np.stack([REG * rng.uniform(0.5, 2.0) + rng.uniform(-3, 3, 3) for _ in range(E)])


def _basis(nm, nq):
    b = O.basis_prism(nm, nq[0], nq[2])
    if nq[1] != nq[0]:  # direction 1 on its own GLL points
        a = O.basis(nm, nq[1])
        b.update(nq1=nq[1], Q=nq[0] * nq[1] * nq[2], z1=a["z"], w1=a["w"], D1=a["D"],
                 bdata1=a["bdata"], dbdata1=a["dbdata"])
    return b


def _tables(b):
    return [dict(bdata=b[f"bdata{d}"], dbdata=b[f"dbdata{d}"], D=b[f"D{d}"],
                 w=b["w2_raw"] if d == 2 else b[f"w{d}"]) for d in range(3)]


def _rel(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


@pytest.mark.parametrize("nm,nq", COMBOS)
@pytest.mark.parametrize("deformed", [False, True])
def test_prism_operators_match_the_oracle(nm, nq, deformed):
    print(f"[case] prism nm={nm} nq={nq} deformed={deformed}", flush=True)
    E = 301
    b = _basis(nm, nq)
    verts = _mesh(E, nm * 100 + nq[0], deformed)
    jac, df = O.geom_prism(verts, b, deformed)
    gj, gd = H.geom_device("prism", torch.from_numpy(verts).cuda(), nq, None, deformed)
    assert np.abs(gj.cpu().numpy() - jac).max() < 1e-13 * np.abs(jac).max()
    assert np.abs(gd.cpu().numpy() - df).max() < 1e-13 * np.abs(df).max()
    M, Q = b["M"], b["Q"]
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    geo = "Deformed" if deformed else "Regular"
    tabs = _tables(b)
    du, dfq = torch.from_numpy(u).cuda(), torch.from_numpy(f).cuda()

    bw = H.Operator("BwdTrans_Prism_Regular", nm, nq, E, basis=tabs)
    assert _rel(bw.apply(du).cpu().numpy(), O.bwdtrans_prism(b, E, u)) < 1e-12

    ip = H.Operator(f"IProduct_Prism_{geo}", nm, nq, E, basis=tabs)
    ip.set_jac(gj)
    assert _rel(ip.apply(dfq).cpu().numpy(), O.iproduct_prism(b, E, deformed, jac, f)) < 1e-12

    hm = H.Operator(f"Helmholtz_Prism_{geo}", nm, nq, E, basis=tabs)
    hm.set_jac(gj)
    hm.set_df(gd)
    hm.set_lambda(2.5)
    got = hm.apply(du).cpu().numpy()
    assert _rel(got, O.helmholtz_prism(b, E, deformed, jac, df, 2.5, u)) < 1e-12

    pd = H.Operator(f"PhysDeriv_Prism_{geo}", nm, nq, E, basis=tabs)
    pd.set_df(gd)
    g = pd.apply(dfq).cpu().numpy()
    ref = O.physderiv_prism(b, E, deformed, df, f)
    for c in range(3):
        assert _rel(g[c * E * Q:(c + 1) * E * Q], ref[c]) < 1e-12

    ipd = H.Operator(f"IProductWRTDerivBase_Prism_{geo}", nm, nq, E, basis=tabs)
    ipd.set_jac(gj)
    ipd.set_df(gd)
    fs = [rng.uniform(-1, 1, E * Q) for _ in range(3)]
    dfs = torch.from_numpy(np.concatenate(fs)).cuda()
    got = ipd.apply(dfs).cpu().numpy()
    assert _rel(got, O.iprodderiv_prism(b, E, deformed, jac, df, fs)) < 1e-12
    with pytest.raises(H.MFError):
        ipd.apply_transpose(du)

    # Euclidean transposes: <A x, y> = <x, A^T y>
    for op, nin, nout in [(bw, M, Q), (ip, Q, M), (hm, M, M)]:
        x = torch.from_numpy(rng.uniform(-1, 1, E * nin)).cuda()
        y = torch.from_numpy(rng.uniform(-1, 1, E * nout)).cuda()
        lhs = torch.dot(op.apply(x), y).item()
        rhs = torch.dot(x, op.apply_transpose(y)).item()
        assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_prism_many_elements_per_workgroup_loop():
    """More elements than workgroups: each workgroup loops over several elements."""
    print("[case] prism E=20000", flush=True)
    nm, nq, E = 3, (4, 4, 3), 20000
    b = _basis(nm, nq)
    verts = _mesh(E, 5, True)
    jac, df = O.geom_prism(verts, b, True)
    u = np.random.default_rng(2).uniform(-1, 1, E * b["M"])
    hm = H.Operator("Helmholtz_Prism_Deformed", nm, nq, E)  # library-generated tables
    hm.set_jac(jac)
    hm.set_df(df)
    hm.set_lambda(0.5)
    got = hm.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert _rel(got, O.helmholtz_prism(b, E, True, jac, df, 0.5, u)) < 1e-12


def _sol(x, y, z, diff):
    s = np.zeros_like(x)
    for k in range(6):
        for l in range(6):
            for m in range(6 - k):
                if not diff:
                    s += x ** k * y ** l * z ** m
                else:
                    s += (k * x ** max(k - 1, 0) * y ** l * z ** m
                          + l * x ** k * y ** max(l - 1, 0) * z ** m
                          + m * x ** k * y ** l * z ** max(m - 1, 0))
    return s


@pytest.mark.parametrize("diff,l2,linf,tol", [(False, 5.16422e-12, 2.08524e-11, 1e-8),
                                              (True, 2.41945e-11, 1.03926e-10, 1e-9)])
def test_prism_locproject_golden_on_the_gpu(diff, l2, linf, tol):
    """LocProject3D_Prism_Mod_Basis_P6_Q7.tst and LocProject_Diff3D_Reg_Prism_Mod_Basis_P6_Q7
    .tst with the GPU BwdTrans / IProduct (and PhysDeriv) forming the projection."""
    print(f"[case] prism locproject diff={diff}", flush=True)
    b = O.basis_prism(6, 7, 7)
    nq = (7, 7, 7)
    tabs = _tables(b)
    v = REG[None]
    gj, gd = H.geom_device("prism", torch.from_numpy(v).cuda(), nq, None, False)
    x, y, z = O.coords_prism(v, b)
    sol = _sol(x, y, z, False)
    if diff:
        pd = H.Operator("PhysDeriv_Prism_Regular", 6, nq, 1, basis=tabs)
        pd.set_df(gd)
        g = pd.apply(torch.from_numpy(sol).cuda()).cpu().numpy()
        sol = g.reshape(3, -1).sum(0)
    bw = H.Operator("BwdTrans_Prism_Regular", 6, nq, 1, basis=tabs)
    ip = H.Operator("IProduct_Prism_Regular", 6, nq, 1, basis=tabs)
    ip.set_jac(gj)
    M = b["M"]
    eye = torch.eye(M, dtype=torch.float64).cuda()
    mass = torch.stack([ip.apply(bw.apply(eye[i].contiguous())) for i in range(M)], 1)
    coef = np.linalg.solve(mass.cpu().numpy(), ip.apply(torch.from_numpy(sol).cuda()).cpu().numpy())
    phys = bw.apply(torch.from_numpy(coef).cuda()).cpu().numpy()
    exact = _sol(x, y, z, True) if diff else sol
    Wq = np.einsum("k,j,i->kji", b["w2"], b["w1"], b["w0"]).ravel()
    jac = gj.cpu().numpy()[0]
    assert abs(np.sqrt(np.sum(Wq * jac * (phys - exact) ** 2)) - l2) <= tol
    assert abs(np.abs(phys - exact).max() - linf) <= tol


def test_prism_rejects_what_it_does_not_cover():
    tabs = _tables(O.basis_prism(4, 5))
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Prism_Regular", 4, 5, 4, precision=H.FP32, basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("IProductWRTDerivBase_Prism_Regular", 4, 5, 4, precision=H.FP32,
                   basis=tabs)
    with pytest.raises(H.MFError):
        H.Operator("Helmholtz_Prism_Regular", 9, 10, 4)
    with pytest.raises(H.MFError):  # one element's arrays exceed a workgroup's LDS
        H.Operator("Helmholtz_Prism_Regular", 8, (16, 16, 16), 4)
    op = H.Operator("PhysDeriv_Prism_Regular", 4, 5, 4, basis=tabs)
    op.set_df(np.zeros(36))
    with pytest.raises(H.MFError):
        op.apply_transpose(torch.zeros(3 * 4 * 100, dtype=torch.float64).cuda())
