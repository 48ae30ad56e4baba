"""GPU parity of the matrix-free operators against the CPU oracle, through the C ABI.

Bar: FP64 outputs within 1e-12 relative L2 of the oracle (the reference CPU path);
FP32 variants within 2e-6 relative L2 (single-precision storage and arithmetic).
"""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O
from ithaca_hip import mesh

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6

HEX_COMBOS = [(2, 2), (2, 4), (3, 5), (4, 4), (4, 8), (5, 6), (6, 7), (6, 12), (7, 9),
              (8, 8), (8, 16)]
QUAD_COMBOS = [(2, 3), (3, 6), (5, 6), (6, 7), (7, 14), (8, 16)]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(shape, nm, nq, deformed, seed, n=3):
    print(f"[case] {shape} nm={nm} nq={nq} deformed={deformed} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    if shape == "hex":
        verts = mesh.hex_box(n, 0.1 if deformed else 0.0, rng)
        geom = O.geom_hex
    else:
        verts = mesh.quad_box(n + 2, 0.1 if deformed else 0.0, rng)
        geom = O.geom_quad
    b = O.basis(nm, nq)
    jac, df = geom(verts, nq, b["z"], deformed)
    return rng, verts, b, jac, df


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype)


def _op(key, nm, nq, E, b, jac=None, df=None, prec=H.FP64):
    op = H.Operator(key, nm, nq, E, precision=prec)
    if jac is not None:
        op.set_jac(jac)
    if df is not None:
        op.set_df(df)
    return op


@pytest.mark.parametrize("shape,nm,nq", [("hex", c[0], c[1]) for c in HEX_COMBOS] +
                         [("quad", c[0], c[1]) for c in QUAD_COMBOS])
def test_bwdtrans_parity(shape, nm, nq):
    rng, verts, b, jac, df = _case(shape, nm, nq, False, 1)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    u = rng.uniform(-1, 1, E * nm ** dim)
    ref = O.bwdtrans(shape, b, E, u)
    op = _op(f"BwdTrans_{shape.capitalize()}_Regular", nm, nq, E, b)
    got = op.apply(_dev(u)).cpu().numpy()
    assert rel(got, ref) < TOL64
    op32 = _op(f"BwdTrans_{shape.capitalize()}_Regular", nm, nq, E, b, prec=H.FP32)
    got32 = op32.apply(_dev(u, torch.float32)).double().cpu().numpy()
    assert rel(got32, ref) < TOL32


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("shape,nm,nq", [("hex", c[0], c[1]) for c in HEX_COMBOS] +
                         [("quad", c[0], c[1]) for c in QUAD_COMBOS])
def test_iproduct_parity(shape, nm, nq, deformed):
    rng, verts, b, jac, df = _case(shape, nm, nq, deformed, 2)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    f = rng.uniform(-1, 1, E * nq ** dim)
    ref = O.iproduct(shape, b, E, deformed, jac, f)
    key = f"IProduct_{shape.capitalize()}_{'Deformed' if deformed else 'Regular'}"
    got = _op(key, nm, nq, E, b, jac).apply(_dev(f)).cpu().numpy()
    assert rel(got, ref) < TOL64
    got32 = _op(key, nm, nq, E, b, jac, prec=H.FP32).apply(_dev(f, torch.float32))
    assert rel(got32.double().cpu().numpy(), ref) < TOL32


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("shape,nm,nq", [("hex", c[0], c[1]) for c in HEX_COMBOS] +
                         [("quad", c[0], c[1]) for c in QUAD_COMBOS])
def test_helmholtz_parity(shape, nm, nq, deformed):
    rng, verts, b, jac, df = _case(shape, nm, nq, deformed, 3)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    u = rng.uniform(-1, 1, E * nm ** dim)
    lam = 2.5
    ref = O.helmholtz(shape, b, E, deformed, jac, df, lam, u)
    key = f"Helmholtz_{shape.capitalize()}_{'Deformed' if deformed else 'Regular'}"
    op = _op(key, nm, nq, E, b, jac, df)
    op.set_lambda(lam)
    got = op.apply(_dev(u)).cpu().numpy()
    assert rel(got, ref) < TOL64
    op32 = _op(key, nm, nq, E, b, jac, df, prec=H.FP32)
    op32.set_lambda(lam)
    got32 = op32.apply(_dev(u, torch.float32)).double().cpu().numpy()
    assert rel(got32, ref) < TOL32 * 10


@pytest.mark.parametrize("shape", ["hex", "quad"])
def test_geometry_kernel_parity(shape):
    nq = 7
    rng = np.random.default_rng(4)
    verts = mesh.hex_box(4, 0.1, rng) if shape == "hex" else mesh.quad_box(6, 0.1, rng)
    b = O.basis(6, nq)
    for deformed in (True, False):
        geom = O.geom_hex if shape == "hex" else O.geom_quad
        jac, df = geom(verts, nq, b["z"], deformed)
        gj, gd = H.geom_device(shape, _dev(verts), nq, b["z"], deformed)
        torch.cuda.synchronize()
        assert rel(gj.cpu().numpy(), jac) < 1e-14
        assert rel(gd.cpu().numpy(), df) < 1e-14


@pytest.mark.parametrize("key", ["BwdTrans_Hex_Regular", "IProduct_Hex_Deformed",
                                 "Helmholtz_Hex_Deformed", "IProduct_Quad_Regular",
                                 "BwdTrans_Quad_Regular", "Helmholtz_Quad_Deformed"])
def test_transpose_is_adjoint(key):
    shape = key.split("_")[1].lower()
    deformed = key.endswith("Deformed")
    nm, nq = (6, 7) if shape == "hex" else (5, 6)
    rng, verts, b, jac, df = _case(shape, nm, nq, deformed, 9)
    E = verts.shape[0]
    op = _op(key, nm, nq, E, b, jac if not key.startswith("BwdTrans") else None,
             df if key.startswith("Helmholtz") else None)
    u = _dev(rng.uniform(-1, 1, op.in_size))
    v = _dev(rng.uniform(-1, 1, op.out_size))
    lhs = torch.dot(v, op.apply(u)).item()
    rhs = torch.dot(op.apply_transpose(v), u).item()
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_host_apply_matches_device_apply():
    rng, verts, b, jac, df = _case("hex", 6, 7, True, 5)
    E = verts.shape[0]
    op = _op("Helmholtz_Hex_Deformed", 6, 7, E, b, jac, df)
    op.set_lambda(10.0)
    u = rng.uniform(-1, 1, E * 216)
    dev = op.apply(_dev(u)).cpu().numpy()
    host = op.apply_host(u)
    assert np.array_equal(dev, host)
    assert op.ndof() == E * 216
    assert op.needs_jac() and op.needs_df()


@pytest.mark.parametrize("key,shape,nm,nq,deformed", [
    ("Helmholtz_Hex_Deformed", "hex", 6, 7, True), ("Helmholtz_Hex_Regular", "hex", 4, 5, False),
    ("BwdTrans_Hex_Regular", "hex", 5, 6, True), ("IProduct_Hex_Deformed", "hex", 6, 7, True),
    ("Mass_Hex_Deformed", "hex", 3, 4, True), ("Helmholtz_Quad_Deformed", "quad", 5, 6, True)])
def test_pipelined_host_apply_is_bitwise_the_device_apply(key, shape, nm, nq, deformed):
    """The host entry runs in element chunks through pinned buffers; every element count
    (fewer elements than chunks, a ragged last chunk, a larger operator after a smaller
    one reusing the staging buffers) gives the device apply's result bit for bit."""
    rng, verts, b, jac, df = _case(shape, nm, nq, deformed, 21, n=11 if shape == "hex" else 30)
    dim = 3 if shape == "hex" else 2
    Q = nq ** dim
    nj = Q if deformed else 1
    for E in (1, 5, 1001, verts.shape[0]):
        jE = np.ascontiguousarray(jac[:E * nj])
        dE = np.ascontiguousarray(df.reshape(dim * dim, -1)[:, :E * nj])
        op = _op(key, nm, nq, E, b, jE if not key.startswith("BwdTrans") else None,
                 dE if key.startswith("Helmholtz") else None)
        if key.startswith("Helmholtz"):
            op.set_lambda(3.0)
        for rep in range(2):
            u = rng.uniform(-1, 1, op.in_size)
            dev = op.apply(_dev(u)).cpu().numpy()
            assert np.array_equal(op.apply_host(u), dev), (E, rep)


def test_error_paths():
    b = O.basis(6, 7)
    with pytest.raises(H.MFError, match="No such operator key"):
        H.Operator("Helmholtz_Cube_Deformed", 6, 7, 4)
    with pytest.raises(H.MFError, match="No such operator key"):
        H.Operator("BwdTrans_Hex_Deformed", 6, 7, 4)
    with pytest.raises(H.MFError, match="modes / points combo"):
        H.Operator("Helmholtz_Hex_Deformed", 6, 13, 4)
    op = H.Operator("Helmholtz_Hex_Deformed", 6, 7, 4)
    with pytest.raises(H.MFError, match="not set"):
        op.apply(torch.zeros(4 * 216, dtype=torch.float64, device="cuda"))
    with pytest.raises(H.MFError, match="expected"):
        op.set_jac(np.zeros(10))
    del b


def test_full_hex_workload_parity():
    """48^3 jittered hexes, P=5, 7 GLL points, lambda=10: every element checked."""
    w = mesh.workload_hex_helmholtz()
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    print(f"[case] full hex workload E={E}", flush=True)
    b = O.basis(nm, nq)
    O.set_threads(16)
    jac, df = O.geom_hex(w["verts"], nq, b["z"], True)
    ref = O.helmholtz("hex", b, E, True, jac, df, w["lam"], w["u"])
    gj, gd = H.geom_device("hex", _dev(w["verts"]), nq, b["z"], True)
    op = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E)
    op.set_jac(gj)
    op.set_df(gd)
    op.set_lambda(w["lam"])
    u = _dev(w["u"])
    got = op.apply(u).cpu().numpy()
    assert rel(got, ref) < TOL64
    # linearity and symmetry at full size
    v = torch.rand_like(u) * 2 - 1
    Hu = op.apply(u)
    Hv = op.apply(v)
    H2 = op.apply(2.0 * u - 3.0 * v)
    assert (torch.linalg.norm(H2 - (2.0 * Hu - 3.0 * Hv)) / torch.linalg.norm(H2)).item() < 1e-13
    a, c = torch.dot(v, Hu).item(), torch.dot(u, Hv).item()
    assert abs(a - c) < 1e-12 * abs(a)
    # FP32 variant of the same workload
    op32 = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, precision=H.FP32)
    op32.set_jac(gj)
    op32.set_df(gd)
    op32.set_lambda(w["lam"])
    got32 = op32.apply(u.float()).double().cpu().numpy()
    assert rel(got32, ref) < 1e-5


def test_quad_workload_parity():
    """32x32 regular quads, P=4, 6 GLL points, lambda=1, seed 0."""
    w = mesh.workload_quad_helmholtz()
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    b = O.basis(nm, nq)
    jac, df = O.geom_quad(w["verts"], nq, b["z"], False)
    ref = O.helmholtz("quad", b, E, False, jac, df, w["lam"], w["u"])
    op = H.Operator("Helmholtz_Quad_Regular", nm, nq, E)
    op.set_jac(jac)
    op.set_df(df)
    op.set_lambda(w["lam"])
    got = op.apply(_dev(w["u"])).cpu().numpy()
    assert rel(got, ref) < TOL64


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("nm,nq", [(2, 3), (4, 8), (6, 7), (8, 9)])
def test_mass_parity(nm, nq, deformed):
    """Fused Mass_Hex_* (MassMatrixOp_MatFree: BwdTrans then IProductWRTBase) against the
    oracle's composition, for one vector and for several vectors in one launch with padded
    rows (leading dimension > vector size); Mass^T == Mass."""
    rng, verts, b, jac, df = _case("hex", nm, nq, deformed, 7)
    E = verts.shape[0]
    M = nm ** 3
    key = f"Mass_Hex_{'Deformed' if deformed else 'Regular'}"
    op = _op(key, nm, nq, E, b, jac=jac)
    assert op.needs_jac() and not op.needs_df()
    nv, ld = 3, E * M + 5
    U = rng.uniform(-1, 1, (nv, ld))
    refs = [O.iproduct("hex", b, E, deformed, jac, O.bwdtrans("hex", b, E, U[v, :E * M]))
            for v in range(nv)]
    got1 = op.apply(_dev(U[0, :E * M])).cpu().numpy()
    assert rel(got1, refs[0]) < TOL64
    out = torch.full((nv, ld), 7.0, dtype=torch.float64, device="cuda")
    op.apply_many(_dev(U), out)
    o = out.cpu().numpy()
    for v in range(nv):
        assert rel(o[v, :E * M], refs[v]) < TOL64
        assert np.all(o[v, E * M:] == 7.0)  # padding untouched
    gt = op.apply_transpose(_dev(U[0, :E * M])).cpu().numpy()
    assert np.array_equal(gt, got1)
    op32 = _op(key, nm, nq, E, b, jac=jac, prec=H.FP32)
    got32 = op32.apply(_dev(U[0, :E * M], torch.float32)).double().cpu().numpy()
    assert rel(got32, refs[0]) < TOL32


def test_mass_quad_rejected_and_apply_many_loops():
    with pytest.raises(H.MFError):
        H.Operator("Mass_Quad_Regular", 4, 5, 4)
    rng, verts, b, jac, df = _case("hex", 4, 5, True, 9)
    E = verts.shape[0]
    op = _op("Helmholtz_Hex_Deformed", 4, 5, E, b, jac=jac, df=df)
    op.set_lambda(2.0)
    U = _dev(rng.uniform(-1, 1, (2, E * 64)))
    many = op.apply_many(U)
    for v in range(2):
        assert torch.equal(many[v], op.apply(U[v].contiguous()))
    # leading dimensions shorter than the vectors are rejected before any launch
    assert H.lib().imf_apply_many_device(op._h, 1, None, 3, None, 3, None) != 0
    assert "leading dimension" in H.lib().imf_last_error().decode()


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("shape,nq", [("hex", q) for q in (2, 3, 5, 7, 9, 12, 16)] +
                         [("quad", q) for q in (2, 4, 6, 7, 11, 16)])
def test_physderiv_parity(shape, nq, deformed):
    """PhysDeriv_{Hex,Quad}_* (points only, nq 2..16) against the oracle: FP64 1e-12, FP32
    2e-6 relative over all dim outputs."""
    nm = min(nq, 8) if nq >= 2 else 2
    rng, verts, b, jac, df = _case(shape, min(nm, nq), nq, deformed, 3)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    f = rng.uniform(-1, 1, E * nq ** dim)
    ref = np.concatenate(O.physderiv(shape, b, E, deformed, df, f))
    key = f"PhysDeriv_{shape.capitalize()}_{'Deformed' if deformed else 'Regular'}"
    op = _op(key, min(nm, nq), nq, E, b, df=df)
    assert op.needs_df() and not op.needs_jac()
    assert op.in_size == E * nq ** dim and op.out_size == dim * E * nq ** dim
    got = op.apply(_dev(f)).cpu().numpy()
    assert rel(got, ref) < TOL64
    op32 = _op(key, min(nm, nq), nq, E, b, df=df, prec=H.FP32)
    got32 = op32.apply(_dev(f, torch.float32)).double().cpu().numpy()
    assert rel(got32, ref) < TOL32
    with pytest.raises(H.MFError):
        op.apply_transpose(_dev(ref))


@pytest.mark.parametrize("deformed", [True, False])
@pytest.mark.parametrize("shape,nm,nq", [("hex", c[0], c[1]) for c in HEX_COMBOS] +
                         [("quad", c[0], c[1]) for c in QUAD_COMBOS])
def test_iprod_deriv_parity(shape, nm, nq, deformed):
    """IProductWRTDerivBase_{Hex,Quad}_* against the oracle (the reference's three
    IProducts with dbdata): FP64 1e-12, FP32 2e-5 relative."""
    rng, verts, b, jac, df = _case(shape, nm, nq, deformed, 4)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    Q = nq ** dim
    ins = [rng.uniform(-1, 1, E * Q) for _ in range(dim)]
    ref = O.iprod_deriv(shape, b, E, deformed, jac, df, ins)
    key = f"IProductWRTDerivBase_{shape.capitalize()}_{'Deformed' if deformed else 'Regular'}"
    op = _op(key, nm, nq, E, b, jac=jac, df=df)
    assert op.needs_jac() and op.needs_df()
    assert op.in_size == dim * E * Q and op.out_size == E * nm ** dim
    x = _dev(np.concatenate(ins))
    got = op.apply(x).cpu().numpy()
    assert rel(got, ref) < TOL64
    op32 = _op(key, nm, nq, E, b, jac=jac, df=df, prec=H.FP32)
    got32 = op32.apply(x.float()).double().cpu().numpy()
    assert rel(got32, ref) < 2e-5


@pytest.mark.parametrize("shape,nm,nq", [("hex", 6, 7), ("quad", 5, 6)])
def test_iprod_deriv_is_weighted_transpose_of_physderiv(shape, nm, nq):
    """<IPDB(v), u_hat> = sum_c <v_c, w J PhysDeriv(BwdTrans(u_hat))_c> on the GPU."""
    rng, verts, b, jac, df = _case(shape, nm, nq, True, 5)
    E = verts.shape[0]
    dim = 3 if shape == "hex" else 2
    M, Q = nm ** dim, nq ** dim
    S = shape.capitalize()
    bwd = _op(f"BwdTrans_{S}_Regular", nm, nq, E, b)
    pd = _op(f"PhysDeriv_{S}_Deformed", nm, nq, E, b, df=df)
    ipd = _op(f"IProductWRTDerivBase_{S}_Deformed", nm, nq, E, b, jac=jac, df=df)
    u = _dev(rng.uniform(-1, 1, E * M))
    v = _dev(rng.uniform(-1, 1, dim * E * Q))
    w1 = np.asarray(b["w"])
    wq = np.einsum("k,j,i->kji", w1, w1, w1).ravel() if dim == 3 else np.outer(w1, w1).ravel()
    wj = _dev(np.tile(wq, E) * jac)
    lhs = torch.dot(ipd.apply(v), u).item()
    g = pd.apply(bwd.apply(u)).view(dim, E * Q)
    rhs = sum(torch.dot(v.view(dim, E * Q)[c], wj * g[c]).item() for c in range(dim))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("shape,nm,nq", [("hex", 6, 7), ("hex", 3, 5), ("quad", 5, 6),
                                         ("quad", 8, 16)])
def test_element_counts(shape, nm, nq):
    """One element, a few, and counts that leave the last workgroup partly filled: every
    operator equals the oracle on the first E elements of a deformed mesh."""
    rng = np.random.default_rng(21)
    full = mesh.hex_box(6, 0.1, rng) if shape == "hex" else mesh.quad_box(12, 0.1, rng)
    geom = O.geom_hex if shape == "hex" else O.geom_quad
    b = O.basis(nm, nq)
    dim = 3 if shape == "hex" else 2
    S = shape.capitalize()
    for E in (1, 2, 3, 5, 17, 130):
        print(f"[case] {shape} nm={nm} nq={nq} nElmt={E}", flush=True)
        verts = np.ascontiguousarray(full[:E])
        jac, df = geom(verts, nq, b["z"], True)
        u = rng.uniform(-1, 1, E * nm ** dim)
        f = rng.uniform(-1, 1, E * nq ** dim)
        ref_b = O.bwdtrans(shape, b, E, u)
        ref_i = O.iproduct(shape, b, E, True, jac, f)
        ref_h = O.helmholtz(shape, b, E, True, jac, df, 1.5, u)
        for prec, tol, dt in ((H.FP64, TOL64, torch.float64), (H.FP32, TOL32 * 10, torch.float32)):
            got = _op(f"BwdTrans_{S}_Regular", nm, nq, E, b, prec=prec).apply(_dev(u, dt))
            assert rel(got.double().cpu().numpy(), ref_b) < tol
            got = _op(f"IProduct_{S}_Deformed", nm, nq, E, b, jac, prec=prec).apply(_dev(f, dt))
            assert rel(got.double().cpu().numpy(), ref_i) < tol
            op = _op(f"Helmholtz_{S}_Deformed", nm, nq, E, b, jac, df, prec=prec)
            op.set_lambda(1.5)
            assert rel(op.apply(_dev(u, dt)).double().cpu().numpy(), ref_h) < tol


@pytest.mark.parametrize("nq", [5, 8])
def test_reference_collection_lambda_zero_cases(nq):
    """library/UnitTests/Collections/TestHexCollection.cpp:3649-3745 and :3747-3845
    (TestHexHelmholtz_MatrixFree_UniformP and ..._Deformed_OverInt): ten copies of the hex
    with vertex 6 moved to (2, 3, 4), Modified_A with 4 modes on 5 (and, over-integrated, 8)
    GLL points, lambda = 0, every coefficient 1; the MatrixFree result must match the
    elemental Helmholtz matrix to BOOST_CHECK_CLOSE's 1e-8 percent after values below 1e-14
    are set to zero.  Here the elemental matrix is the dense one of tests/dense_ref.py."""
    import dense_ref as DR
    print(f"[case] reference lambda = 0 hex, nq={nq}", flush=True)
    v = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [2, 3, 4], [-1, 1, 1]], dtype=float)
    nm, E = 4, 10
    verts = np.repeat(v[None], E, 0)
    b = H.basis_modified_a(nm, nq)
    jac, df = H.geom_device("hex", torch.from_numpy(verts).cuda(), nq, b["z"], True)
    op = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    op.set_jac(jac)
    op.set_df(df)
    op.set_lambda(0.0)
    got = op.apply(torch.ones(E * nm ** 3, dtype=torch.float64, device="cuda")).cpu().numpy()
    m = DR.element_mats("hex", nm, nq, v)
    ref = np.tile(DR.helmholtz_matrix(m, 0.0) @ np.ones(nm ** 3), E)
    got[np.abs(got) < 1e-14] = 0.0
    ref[np.abs(ref) < 1e-14] = 0.0
    close = np.abs(got - ref) <= 1e-10 * np.minimum(np.abs(got), np.abs(ref))  # the strong form
    assert np.all(close), np.abs(got - ref).max()
