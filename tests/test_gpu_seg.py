"""Segment operators on the GPU (FP64, ncoord = 1, 2, 3) against the CPU oracle
(oracle/mf_oracle_seg.c, pinned by closed forms in tests/test_seg_oracle.py), and the
reference's MatrixFree segment unit-test cases
(library/UnitTests/Collections/TestSegCollection.cpp) against dense per-element matrices.
Bars: 1e-12 relative against the oracle; the reference cases' BOOST_CHECK_CLOSE tolerance of
1e-8 percent (1e-10 relative) per entry."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


def _mesh(E, ncoord, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-2, 2, (E, 2, ncoord))
    v[:, 1] += np.where(np.abs(v[:, 1] - v[:, 0]) < 0.1, 0.5, 0.0)  # no short segments
    return v


@pytest.mark.parametrize("nm,nq", [(4, 5), (4, 6), (2, 2), (7, 9), (16, 20), (8, 32)])
@pytest.mark.parametrize("ncoord", [1, 2, 3])
@pytest.mark.parametrize("deformed", [False, True])
def test_seg_operators_match_the_oracle(nm, nq, ncoord, deformed):
    print(f"[case] seg nm={nm} nq={nq} ncoord={ncoord} deformed={deformed}", flush=True)
    E = 1001
    b = O.basis(nm, nq)
    verts = _mesh(E, ncoord, nm * 100 + nq + ncoord)
    jac, df = O.geom_seg(verts, nq, deformed)
    gj, gd = H.geom_device("seg", torch.from_numpy(verts).cuda(), nq, None, deformed)
    assert np.abs(gj.cpu().numpy() - jac).max() < 1e-14 * np.abs(jac).max()
    assert np.abs(gd.cpu().numpy() - df).max() < 1e-14 * np.abs(df).max()
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, E * nm)
    f = rng.uniform(-1, 1, E * nq)
    fs = rng.uniform(-1, 1, ncoord * E * nq)
    du, dfq = torch.from_numpy(u).cuda(), torch.from_numpy(f).cuda()
    geo = "Deformed" if deformed else "Regular"

    bw = H.Operator("BwdTrans_Seg_Regular", nm, nq, E, basis=b, coord_dim=ncoord)
    assert _rel(bw.apply(du).cpu().numpy(), O.bwdtrans_seg(b, E, u)) < 1e-12

    ip = H.Operator(f"IProduct_Seg_{geo}", nm, nq, E, basis=b, coord_dim=ncoord)
    ip.set_jac(jac)
    assert _rel(ip.apply(dfq).cpu().numpy(), O.iproduct_seg(b, E, deformed, jac, f)) < 1e-12

    pd = H.Operator(f"PhysDeriv_Seg_{geo}", nm, nq, E, basis=b, coord_dim=ncoord)
    pd.set_df(df)
    assert pd.out_size == ncoord * E * nq
    got = pd.apply(dfq).cpu().numpy()
    assert _rel(got, O.physderiv_seg(b, E, deformed, ncoord, df, f)) < 1e-12

    ipd = H.Operator(f"IProductWRTDerivBase_Seg_{geo}", nm, nq, E, basis=b, coord_dim=ncoord)
    ipd.set_jac(jac)
    ipd.set_df(df)
    assert ipd.in_size == ncoord * E * nq
    got = ipd.apply(torch.from_numpy(fs).cuda()).cpu().numpy()
    assert _rel(got, O.iprodderiv_seg(b, E, deformed, ncoord, jac, df, fs)) < 1e-12
    for op, n in [(ipd, E * nm), (pd, ncoord * E * nq)]:
        with pytest.raises(H.MFError):
            op.apply_transpose(torch.zeros(n, dtype=torch.float64).cuda())

    # Euclidean transposes: <A x, y> = <x, A^T y>
    for op, nin, nout in [(bw, nm, nq), (ip, nq, nm)]:
        x = torch.from_numpy(rng.uniform(-1, 1, E * nin)).cuda()
        y = torch.from_numpy(rng.uniform(-1, 1, E * nout)).cuda()
        lhs = torch.dot(op.apply(x), y).item()
        rhs = torch.dot(x, op.apply_transpose(y)).item()
        assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_seg_rejects_what_the_reference_does_not_register():
    b = O.basis(4, 6)
    for key in ["Helmholtz_Seg_Regular", "BwdTrans_Seg_Deformed", "Mass_Seg_Regular"]:
        with pytest.raises(H.MFError):
            H.Operator(key, 4, 6, 3, basis=b)
    with pytest.raises(H.MFError):
        H.Operator("IProduct_Seg_Regular", 4, 6, 3, basis=b, precision=H.FP32)
    with pytest.raises(H.MFError):
        H.Operator("IProduct_Seg_Regular", 4, 6, 3, basis=b, coord_dim=4)
    with pytest.raises(H.MFError):
        H.Operator("IProduct_Seg_Regular", 6, 4, 3, basis=O.basis(6, 4))


# ---- the reference's MatrixFree segment cases: nelmts = 10 copies of one segment, Modified_A
# with 4 modes on GLL points, compared per entry with the element-wise StdExpansion result.

def _dense(nm, nq, v0, v1):
    b = O.basis(nm, nq)
    B = b["bdata"].reshape(nm, nq)
    dB = b["dbdata"].reshape(nm, nq)
    D = b["D"].reshape(nq, nq)
    v0, v1 = np.asarray(v0, float), np.asarray(v1, float)
    t = 0.5 * (v1 - v0)
    J = np.linalg.norm(t)
    x = v0[:, None] + np.outer(t, 1 + b["z"])  # (ncoord, nq)
    return b, B, dB, D, t, J, x


def _close(got, ref):
    assert np.all(np.abs(got - ref) <= 1e-10 * np.abs(ref) + 1e-14 * np.abs(ref).max())


def _geom(v0, v1, nq):
    verts = np.array([[v0, v1]] * 10, dtype=float)
    return O.geom_seg(verts, nq, False)


def test_reference_seg_bwdtrans_matrixfree_case():
    """TestSegCollection.cpp:283 TestSegBwdTrans_MatrixFree_UniformP_MultiElmt (coefficients
    all one, 6 points)."""
    print("[case] reference seg bwdtrans", flush=True)
    b, B, *_ = _dense(4, 6, [-1.0], [1.0])
    op = H.Operator("BwdTrans_Seg_Regular", 4, 6, 10, basis=b)
    got = op.apply(torch.ones(40, dtype=torch.float64).cuda()).cpu().numpy()
    _close(got, np.tile(B.T @ np.ones(4), 10))


def test_reference_seg_iproduct_matrixfree_case():
    """TestSegCollection.cpp:512 TestSegIProductWRTBase_MatrixFree_UniformP_MultiElmt
    (-1.5 to 1, 5 points, phys = sin(x))."""
    print("[case] reference seg iproduct", flush=True)
    b, B, dB, D, t, J, x = _dense(4, 5, [-1.5], [1.0])
    jac, _ = _geom([-1.5], [1.0], 5)
    op = H.Operator("IProduct_Seg_Regular", 4, 5, 10, basis=b)
    op.set_jac(jac)
    phys = np.sin(x[0])
    got = op.apply(torch.from_numpy(np.tile(phys, 10)).cuda()).cpu().numpy()
    _close(got, np.tile(B @ (b["w"] * J * phys), 10))


@pytest.mark.parametrize("ncoord,v1,line", [(1, [1.0], 861), (2, [1.0, 1.0], 924),
                                            (3, [1.0, 1.0, 1.0], 990)])
def test_reference_seg_physderiv_matrixfree_cases(ncoord, v1, line):
    """TestSegCollection.cpp:861 / :924 / :990 TestSegPhysDeriv_MatrixFree_UniformP_MultiElmt
    _1D / _2D / _3D: (-1.5, 0, 0) to v1, 5 points, phys = sin x, sin x cos y,
    sin x cos y sin z."""
    print(f"[case] reference seg physderiv {ncoord}D (line {line})", flush=True)
    v0 = [-1.5] + [0.0] * (ncoord - 1)
    b, B, dB, D, t, J, x = _dense(4, 5, v0, v1)
    _, df = _geom(v0, v1, 5)
    phys = np.sin(x[0])
    if ncoord > 1:
        phys = phys * np.cos(x[1])
    if ncoord > 2:
        phys = phys * np.sin(x[2])
    op = H.Operator("PhysDeriv_Seg_Regular", 4, 5, 10, basis=b, coord_dim=ncoord)
    op.set_df(df)
    got = op.apply(torch.from_numpy(np.tile(phys, 10)).cuda()).cpu().numpy()
    dxi = D.T @ phys
    ref = np.concatenate([np.tile(t[c] / (t @ t) * dxi, 10) for c in range(ncoord)])
    _close(got, ref)


@pytest.mark.parametrize("ncoord", [1, 2])
def test_reference_seg_iprodderiv_matrixfree_cases(ncoord):
    """TestSegCollection.cpp:1493 TestSegIProductWRTDerivBase_MatrixFree_UniformP_MultiElmt
    ((-1) to (1), phys = sin x) and :1559 _CoordimTwo ((-1, 0) to (1, 1), inputs sin x and
    cos y, the two directions' results summed); 6 points."""
    print(f"[case] reference seg iprodderiv coordim {ncoord}", flush=True)
    v0, v1 = ([-1.0], [1.0]) if ncoord == 1 else ([-1.0, 0.0], [1.0, 1.0])
    b, B, dB, D, t, J, x = _dense(4, 6, v0, v1)
    jac, df = _geom(v0, v1, 6)
    ins = [np.sin(x[0])] + ([np.cos(x[1])] if ncoord == 2 else [])
    op = H.Operator("IProductWRTDerivBase_Seg_Regular", 4, 6, 10, basis=b, coord_dim=ncoord)
    op.set_jac(jac)
    op.set_df(df)
    fin = np.concatenate([np.tile(p, 10) for p in ins])
    got = op.apply(torch.from_numpy(fin).cuda()).cpu().numpy()
    ref = sum(dB @ (b["w"] * J * (t[c] / (t @ t)) * ins[c]) for c in range(ncoord))
    _close(got, np.tile(ref, 10))
