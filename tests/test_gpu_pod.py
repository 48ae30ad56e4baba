"""GPU checks of the POD offline kernels against FP64 CPU oracles.

* Gram products (FP64 MFMA) against numpy FP64 to 1e-12 relative; the FP32- and
  BF16-split variants report their error against the FP64 result (bounded here).
* The FP64 Jacobi eigen-decomposition against LAPACK (numpy.linalg.eigh) and the
  reference's Eigen SelfAdjointEigenSolver (oracle/_ref), and by its residual.
* The mass-weighted POD pipeline on a small hex mesh against a dense CPU restatement.
"""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O
import oracle.rom as R
from ithaca_hip import mesh, pod

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("m,n,K", [(512, 512, 200_000), (100, 37, 12_345), (16, 16, 33),
                                   (257, 129, 70_001)])
def test_gram_fp64_parity(m, n, K):
    print(f"[case] gram fp64 m={m} n={n} K={K}", flush=True)
    rng = np.random.default_rng(m + n)
    X = rng.standard_normal((m, K))
    Y = rng.standard_normal((n, K))
    ref = X @ Y.T
    got = pod.gram(_dev(X), _dev(Y)).cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    if m == n:
        refs = X @ X.T
        gs = pod.gram(_dev(X)).cpu().numpy()
        assert np.linalg.norm(gs - refs) / np.linalg.norm(refs) < 1e-12
        assert np.array_equal(gs, gs.T)


@pytest.mark.parametrize("mode,bound", [(pod.GRAM_FP32X3, 1e-6), (pod.GRAM_BF16X3, 1e-4)])
def test_gram_split_precision_error(mode, bound):
    rng = np.random.default_rng(1)
    X = _dev(rng.standard_normal((256, 300_000)))
    ref = pod.gram(X)
    got = pod.gram(X, mode=mode)
    err = (torch.linalg.norm(got - ref) / torch.linalg.norm(ref)).item()
    print(f"mode {mode}: rel err vs FP64 {err:.3e}")
    assert err < bound


@pytest.mark.parametrize("n", [512, 64, 33])
def test_eig_sym_parity(n):
    print(f"[case] eig n={n}", flush=True)
    rng = np.random.default_rng(n)
    # PSD with a wide spectrum and a cluster of tiny eigenvalues (like a POD Gram matrix)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.logspace(4, -8, n // 2), 1e-10 * rng.uniform(0, 1, n - n // 2)])
    G = (U * lam) @ U.T
    G = 0.5 * (G + G.T)
    ev, V, sweeps = pod.eig_sym(_dev(G))
    V = V.cpu().numpy()
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    assert np.abs(ev - ref).max() < 1e-12 * ref[0]
    # residuals and orthogonality
    res = np.linalg.norm(G @ V.T - V.T * ev, axis=0).max()
    assert res < 1e-11 * ref[0]
    assert np.abs(V @ V.T - np.eye(n)).max() < 1e-12
    # leading, well separated eigenvectors against the reference's Eigen solver
    eg = R.eigen_selfadjoint(G)
    if eg is not None:
        evr, Vr = eg
        order = np.argsort(evr)[::-1]
        assert np.abs(ev - evr[order]).max() < 1e-12 * ref[0]
        for j in range(min(8, n)):
            assert abs(abs(V[j] @ Vr[:, order[j]]) - 1.0) < 1e-10
    print(f"sweeps {sweeps}")


def test_eig_sym_two_workgroups_per_block_pair():
    """Order 2100: 66 block pairs, so each pair's rows are split over two workgroups instead
    of four (at most one workgroup per CU); a POD-like spectrum (20 decaying modes over a
    noise floor) against LAPACK eigvalsh.  eig_sym's default tolerance is n eps (pairs stop
    rotating below n eps relative), so the eigenvalues are normwise accurate to a multiple of
    n eps ||G||: the bars are 10 n eps lambda_max for the eigenvalues and 100 n eps
    lambda_max for the residuals (at n = 512 the first is the 1e-12 of the test above);
    the eigenvectors are products of rotations, orthonormal to 1e-12 whatever the order."""
    n = 2100
    print(f"[case] eig n={n}", flush=True)
    rng = np.random.default_rng(21)
    a = rng.standard_normal((n, 20)) * (2.0 ** -np.arange(20))[None, :]
    B = np.concatenate([a, 1e-6 * rng.standard_normal((n, 2 * n))], axis=1)
    G = B @ B.T
    G = 0.5 * (G + G.T)
    ev, V, sweeps = pod.eig_sym(_dev(G))
    V = V.cpu().numpy()
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    neps = n * np.finfo(np.float64).eps
    err = np.abs(ev - ref).max() / ref[0]
    res = np.linalg.norm(G @ V.T - V.T * ev, axis=0).max() / ref[0]
    orth = np.abs(V @ V.T - np.eye(n)).max()
    print(f"sweeps {sweeps}: eigenvalue error {err:.2e}, residual {res:.2e}, "
          f"orthogonality {orth:.2e} (n eps = {neps:.2e})")
    assert err < 10 * neps
    assert res < 100 * neps
    assert orth < 1e-12


def test_pod_pipeline_small_mesh():
    """Mass-weighted POD of 24 snapshots on a 4^3 deformed hex mesh vs a dense FP64
    restatement (oracle mass operator + LAPACK eigh)."""
    rng = np.random.default_rng(2)
    verts = mesh.hex_box(4, 0.1, rng)
    E, nm, nq = verts.shape[0], 6, 7
    N = E * nm ** 3
    ns = 24
    # smooth-ish snapshots: random combinations of a few fixed random fields + noise
    base = rng.standard_normal((6, N))
    S = rng.standard_normal((ns, 6)) @ base * (2.0 ** -np.arange(1, 7))[None, :].sum() + \
        1e-6 * rng.standard_normal((ns, N))
    b = O.basis(nm, nq)
    jac, df = O.geom_hex(verts, nq, b["z"], True)
    MS_ref = np.stack([O.iproduct("hex", b, E, True, jac, O.bwdtrans("hex", b, E, s)) for s in S])
    G_ref = S @ MS_ref.T
    bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E)
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E)
    ipr.set_jac(jac)
    Sd = _dev(S)
    MS = pod.mass_apply(Sd, bwd, ipr)
    G = pod.gram(Sd, MS)
    assert (torch.linalg.norm(G - _dev(G_ref)) / np.linalg.norm(G_ref)).item() < 1e-12
    G = 0.5 * (G + G.T)
    ev, Psi, _ = pod.eig_sym(G)
    evr = np.sort(np.linalg.eigvalsh(0.5 * (G_ref + G_ref.T)))[::-1]
    assert np.abs(ev - evr).max() < 1e-11 * evr[0]
    r = pod.energy_truncation(ev, 1 - 1e-8)
    V = pod.pod_basis(Sd, ev, Psi, r)
    MV = pod.mass_apply(V, bwd, ipr)
    VtMV = pod.gram(V, MV).cpu().numpy()
    assert np.abs(VtMV - np.eye(r)).max() < 1e-9


def test_elemental_projection_vs_dense():
    """Elemental L2 projection (HexExp::v_FwdTrans restated: IProduct then the element's
    inverse mass) by preconditioned CG on the GPU vs a dense per-element solve built from
    independent numpy code (tests/dense_ref.py); the projection of a field already in the
    space returns its coefficients; the GPU quadrature-point coordinates match the dense
    trilinear map."""
    import dense_ref as DR

    rng = np.random.default_rng(5)
    verts = mesh.hex_box(3, 0.1, rng)
    E, nm, nq = verts.shape[0], 6, 7
    M, Q = nm ** 3, nq ** 3
    b = H.basis_modified_a(nm, nq)
    vd = _dev(verts)
    jac, _ = H.geom_device("hex", vd, nq, b["z"], True)
    proj = pod.ElementalProjector(E, nm, nq, b, jac)
    X = mesh.hex_quad_points(vd, b["z"])
    f = (torch.sin(np.pi * X[..., 0]) * torch.sin(2 * np.pi * X[..., 1])
         * torch.cos(3 * np.pi * X[..., 2])).reshape(-1)
    u, it, res = proj.project(f)
    print(f"[case] projection E={E}: {it} CG iterations, residual {res:.2e}", flush=True)
    assert res < 1e-12 and it <= 40
    uh, Xh, fh = u.cpu().numpy().reshape(E, M), X.cpu().numpy(), f.cpu().numpy().reshape(E, Q)
    eps = np.finfo(float).eps
    kmax = 0.0
    for e in range(E):
        m = DR.element_mats("hex", nm, nq, verts[e])
        assert np.abs(m["X"] - Xh[e]).max() < 1e-14
        WJ = m["W"] * m["J"]
        Me = m["B"].T @ (WJ[:, None] * m["B"])
        rhs = m["B"].T @ (WJ * fh[e])
        # residual of the dense system: independent of the conditioning of M_e
        assert np.linalg.norm(Me @ uh[e] - rhs) <= 1e-13 * np.linalg.norm(rhs)
        # forward difference to the dense LAPACK solve: bounded by cond(M_e) eps
        kappa = np.linalg.cond(Me)
        kmax = max(kmax, kappa)
        ue = np.linalg.solve(Me, rhs)
        assert np.linalg.norm(uh[e] - ue) <= 10 * kappa * eps * np.linalg.norm(ue)
    print(f"max cond(M_e) {kmax:.3g}")
    # a field of the space is reproduced (to the same conditioning bound)
    u0 = _dev(rng.uniform(-1, 1, E * M))
    f0 = proj.bwd.apply(u0)
    u1, _, _ = proj.project(f0)
    assert (torch.linalg.norm(u1 - u0) / torch.linalg.norm(u0)).item() < 10 * kmax * eps


@pytest.mark.parametrize("m,K", [(300, 70_001), (512, 131_072), (129, 4_097)])
def test_gram_symmetric_result(m, K):
    """X diag(d) X^T with the symmetric-result flag (upper blocks only, mirrored) against
    numpy FP64, for odd K (8-byte loads) and even K (16-byte loads), all three modes."""
    print(f"[case] gram symmetric-result m={m} K={K}", flush=True)
    rng = np.random.default_rng(m)
    X = rng.standard_normal((m, K))
    d = rng.uniform(0.5, 2.0, K)
    Y = X * d
    ref = X @ Y.T
    Xd, Yd = _dev(X), _dev(Y)
    for mode, tol in ((pod.GRAM_FP64, 1e-12), (pod.GRAM_FP32X3, 1e-6), (pod.GRAM_BF16X3, 1e-4)):
        G = pod.gram(Xd, Yd, mode=mode, symmetric_result=True).cpu().numpy()
        assert np.array_equal(G, G.T)
        assert np.linalg.norm(G - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("nm,nq", [(4, 5), (6, 7), (3, 6)])
def test_galerkin_projection_of_affine_terms(nm, nq):
    """Per-term Galerkin projection C_q = V A_q V^T (CoupledLinearNS_ROM.cpp:2788-2802) of
    the hex Helmholtz terms, stiffness (lambda = 0) and mass, on a deformed 2^3 mesh,
    against dense FP64 element matrices from independent numpy code (tests/dense_ref.py):
    FP64 to 1e-12; the split-precision products report their error against that CPU
    oracle.  The multi-vector apply equals one apply per vector bit for bit, and
    lambda C_mass + C_stiffness equals the projection of the Helmholtz operator."""
    import dense_ref as DR

    print(f"[case] galerkin projection nm={nm} nq={nq}", flush=True)
    rng = np.random.default_rng(9)
    verts = mesh.hex_box(2, 0.1, rng)
    E, M = verts.shape[0], nm ** 3
    N, r = E * M, 7
    b = H.basis_modified_a(nm, nq)
    jac, df = H.geom_device("hex", _dev(verts), nq, b["z"], True)
    terms = pod.helmholtz_affine_terms(nm, nq, E, b, jac, df)
    V = rng.standard_normal((r, N))
    Vd = _dev(V)
    Kref = np.zeros((r, r))
    Mref = np.zeros((r, r))
    for e in range(E):
        m = DR.element_mats("hex", nm, nq, verts[e])
        Ke = DR.helmholtz_matrix(m, 0.0)
        WJ = m["W"] * m["J"]
        Me = m["B"].T @ (WJ[:, None] * m["B"])
        Ve = V[:, e * M:(e + 1) * M]
        Kref += Ve @ Ke @ Ve.T
        Mref += Ve @ Me @ Ve.T
    C = pod.project_affine(terms, Vd)
    for name, ref in (("stiffness", Kref), ("mass", Mref)):
        got = C[name].cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"{name}: FP64 rel err vs dense CPU {err:.2e}")
        assert err < 1e-12
        assert np.array_equal(got, got.T)
    for mode, bound in ((pod.GRAM_FP32X3, 1e-6), (pod.GRAM_BF16X3, 1e-4)):
        Cs = pod.project_affine(terms, Vd, mode=mode)
        for name, ref in (("stiffness", Kref), ("mass", Mref)):
            err = np.linalg.norm(Cs[name].cpu().numpy() - ref) / np.linalg.norm(ref)
            print(f"{name} mode {mode}: rel err vs dense CPU FP64 {err:.2e}")
            assert err < bound
    # multi-vector apply == per-vector applies (bitwise), for both terms
    for op in terms.values():
        many = op.apply_many(Vd)
        for j in range(r):
            assert torch.equal(many[j], op.apply(Vd[j].contiguous()))
    # affine recombination: lambda C_mass + C_stiffness == projection of Helmholtz(lambda)
    lam = 10.0
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    CH, _ = pod.galerkin_project(Hop, Vd)
    comb = lam * C["mass"] + C["stiffness"]
    assert (torch.linalg.norm(CH - comb) / torch.linalg.norm(CH)).item() < 1e-13


def test_galerkin_projection_rejects_non_square_terms():
    b = H.basis_modified_a(4, 5)
    op = H.Operator("BwdTrans_Hex_Regular", 4, 5, 8, basis=b)
    V = torch.zeros((3, op.in_size), dtype=torch.float64, device="cuda")
    with pytest.raises(H.MFError, match="Helmholtz or Mass"):
        pod.galerkin_project(op, V)


def test_method_of_snapshots_matches_reference_bdcsvd():
    """With the identity as weight, Gram + FP64 eigensolve on the GPU reproduce the
    reference's POD (CoupledLinearNS_ROM.cpp:3925-3954: Eigen BDCSVD of the snapshot matrix,
    built from the reference's vendored Eigen 3.3.7 in oracle/_ref): sigma_i^2 = lambda_i,
    and the left singular vectors U_i = S psi_i / sigma_i up to sign, for the modes whose
    sigma is resolved by the method of snapshots (sigma_i / sigma_1 > 1e-3)."""
    if R.eigen_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(21)
    Nrow, k = 3000, 24
    # snapshot matrix with a decaying spectrum: U0 diag(s) W0^T + small noise
    U0, _ = np.linalg.qr(rng.standard_normal((Nrow, k)))
    W0, _ = np.linalg.qr(rng.standard_normal((k, k)))
    s = 10.0 ** -np.linspace(0, 5, k)
    S = (U0 * s) @ W0.T + 1e-9 * rng.standard_normal((Nrow, k))
    sv, Ur = R.eigen_bdcsvd(S)
    St = _dev(S.T)                      # (k, N): snapshots as rows
    G = pod.gram(St)
    ev, Psi, _ = pod.eig_sym(G)
    assert np.abs(ev - sv ** 2).max() <= 1e-10 * sv[0] ** 2
    ok = sv / sv[0] > 1e-3
    Ug = (St.T @ Psi.T).cpu().numpy() / np.sqrt(np.maximum(ev, 1e-300))
    for j in np.nonzero(ok)[0]:
        c = Ug[:, j] @ Ur[:, j]
        assert np.linalg.norm(Ug[:, j] - np.sign(c) * Ur[:, j]) <= 1e-10 * (sv[0] / sv[j]) ** 2, j
    # the reference's RBsize rule on these singular values equals a direct count:
    # RBsize = 1 + #{i : cumulative sigma fraction up to i < tolerance} (capped at k)
    for tol in (0.5, 0.9, 0.99, 0.9999, 0.999999):
        frac = np.cumsum(sv) / sv.sum()
        direct = min(k, 1 + int(np.sum(frac < tol)))
        assert pod.reference_rbsize(sv, tol) == direct
        assert pod.reference_rbsize(np.sqrt(np.maximum(ev, 0.0)), tol) == direct
