"""Hexahedral operators with different modes and points per direction on the GPU (the
runtime-size kernels), against dense element matrices built from independent numpy code
(tests/dense_ref.py), and the reference's solver golden for such an expansion,
Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_Heterogeneous.tst, through the GPU path.
Bar: FP64 relative L2 <= 1e-12 against the dense matrices."""
import os

import numpy as np
import pytest
import torch

import dense_ref as DR
import ithaca_hip as H
from ithaca_hip import assembly, mesh
from ithaca_hip import nektar_io as N

pytestmark = pytest.mark.gpu
COMBOS = [((4, 6, 8), (5, 7, 9)), ((3, 5, 2), (6, 5, 4)), ((2, 7, 4), (3, 12, 4))]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


def _setup(nm3, nq3, deformed, seed=2):
    rng = np.random.default_rng(seed)
    verts = mesh.hex_box(2, 0.1 if deformed else 0.0, rng)
    E = verts.shape[0]
    bases = [H.basis_modified_a(m, q) for m, q in zip(nm3, nq3)]
    jac, df = H.geom_device("hex", _dev(verts), nq3, [b["z"] for b in bases], deformed)
    mats = [DR.element_mats("hex", tuple(nm3), tuple(nq3), verts[e]) for e in range(E)]
    return verts, E, bases, jac, df, mats


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("nm3,nq3", COMBOS)
def test_per_direction_operators_against_dense_matrices(nm3, nq3, deformed):
    print(f"[case] per-direction operators nm={nm3} nq={nq3} deformed={deformed}", flush=True)
    verts, E, bases, jac, df, mats = _setup(nm3, nq3, deformed)
    M, Q = int(np.prod(nm3)), int(np.prod(nq3))
    geo = "Deformed" if deformed else "Regular"
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, E * M)
    f = rng.uniform(-1, 1, E * Q)
    lam = 3.5
    ops = {}
    for kind in ("BwdTrans_Hex_Regular", f"IProduct_Hex_{geo}", f"Helmholtz_Hex_{geo}",
                 f"Mass_Hex_{geo}"):
        op = H.Operator(kind, nm3, nq3, E, basis=bases)
        if op.needs_jac():
            op.set_jac(jac)
        if op.needs_df():
            op.set_df(df)
        op.set_lambda(lam)
        ops[kind.split("_")[0]] = op
    assert ops["BwdTrans"].in_size == E * M and ops["BwdTrans"].out_size == E * Q
    got = {k: op.apply(_dev(f if k == "IProduct" else u)).cpu().numpy() for k, op in ops.items()}
    ref = {k: [] for k in ops}
    for e, m in enumerate(mats):
        ue, fe = u[e * M:(e + 1) * M], f[e * Q:(e + 1) * Q]
        WJ = m["W"] * m["J"]
        ref["BwdTrans"].append(m["B"] @ ue)
        ref["IProduct"].append(m["B"].T @ (WJ * fe))
        ref["Mass"].append(m["B"].T @ (WJ * (m["B"] @ ue)))
        ref["Helmholtz"].append(DR.helmholtz_matrix(m, lam) @ ue)
    for k in ops:
        err = _rel(got[k], np.concatenate(ref[k]))
        print(f"{k}: rel err vs dense {err:.2e}")
        assert err < 1e-12, k
    # several vectors in one launch equal one launch per vector
    U = _dev(rng.uniform(-1, 1, (3, E * M)))
    many = ops["Helmholtz"].apply_many(U)
    for j in range(3):
        assert torch.equal(many[j], ops["Helmholtz"].apply(U[j].contiguous()))


def test_per_direction_operators_reject_what_they_do_not_cover():
    bases = [H.basis_modified_a(m, q) for m, q in zip((4, 6, 8), (5, 7, 9))]
    with pytest.raises(H.MFError, match="FP64"):
        H.Operator("Helmholtz_Hex_Deformed", (4, 6, 8), (5, 7, 9), 8, precision=H.FP32,
                   basis=bases)
    with pytest.raises(H.MFError, match="supported for FP64"):
        H.Operator("PhysDeriv_Hex_Deformed", (4, 6, 8), (5, 7, 9), 8, basis=bases)


def test_helmholtz3d_hex_heterogeneous_golden_on_the_gpu():
    """Helmholtz3D_CG_Hex_Heterogeneous.tst (NUMMODES 4,6,8, NUMPOINTS 5,7,9; L2 error
    5.23385e-05, Linf 0.00014837 at the printed 6 digits, tolerance 1e-12) through the GPU
    path: per-direction geometry, IProduct of the forcing, the per-direction assembly, the
    Dirichlet lift and CG on the GPU, errors at the quadrature points from the GPU
    BwdTrans."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    msh = N.read_mesh(os.path.join(gold, "Helmholtz3D_Hex_Heterogeneous.xml"))
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm3, nq3 = msh.expansions[0]["nummodes"], msh.expansions[0]["numpoints"]
    E, lam = len(ids), 1.0
    print(f"[case] heterogeneous Helmholtz golden nm={nm3} nq={nq3}", flush=True)
    uex = lambda P: np.sin(np.pi / 2 * P[:, 0]) * np.sin(np.pi / 2 * P[:, 1]) * np.sin(np.pi / 2 * P[:, 2])
    bases = [H.basis_modified_a(m, q) for m, q in zip(nm3, nq3)]
    zs = [b["z"] for b in bases]
    jac, df = H.geom_device("hex", _dev(X), nq3, zs, True)
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm3, nq3, E, basis=bases)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator("IProduct_Hex_Deformed", nm3, nq3, E, basis=bases)
    ipr.set_jac(jac)
    bwd = H.Operator("BwdTrans_Hex_Regular", nm3, nq3, E, basis=bases)
    P = assembly._points(X, zs).reshape(-1, 3)
    A = assembly.Assembly(ids, tuple(nm3), dirichlet=True)
    rhs = A.assemble(ipr.apply(_dev((lam + 3 * np.pi ** 2 / 4) * uex(P))))
    g = assembly.dirichlet_lift(X, ids, tuple(nm3), uex, nq=nq3)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-14, max_iter=5000, lift=g)
    d = bwd.apply(A.global_to_local(x)).cpu().numpy() - uex(P)
    W = np.kron(np.kron(bases[2]["w"], bases[1]["w"]), bases[0]["w"])
    wj = jac.cpu().numpy() * np.tile(W, E)
    l2, linf = np.sqrt(np.sum(wj * d * d)), np.abs(d).max()
    print(f"iterations {it}, residual {res:.1e}, L2 {l2:.6e}, Linf {linf:.6e}")
    assert abs(float(f"{l2:.6g}") - 5.23385e-05) <= 1e-12
    assert abs(float(f"{linf:.6g}") - 0.00014837) <= 1e-12
