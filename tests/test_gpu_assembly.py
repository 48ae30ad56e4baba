"""GPU C0 assembly and matrix-free Helmholtz solve against the CPU checks of
tests/test_assembly_cpu.py (the same map, dense element matrices, scipy sparse solve)."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
from ithaca_hip import assembly, mesh
import oracle as O
import os
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import dense_ref as DR
from ithaca_hip import nektar_io as N
from test_assembly_cpu import _mesh, _qmesh, cpu_solve, l2_error, manufactured

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype)


def _gpu_problem(verts, ids, nm, nq, lam, dirichlet=True):
    E = len(ids)
    b = H.basis_modified_a(nm, nq)
    vd = _dev(verts)
    jac, df = H.geom_device("hex", vd, nq, b["z"], True)
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    u, f = manufactured(lam)
    X = mesh.hex_quad_points(vd, b["z"]).reshape(-1, 3)
    fx = ((3 * np.pi ** 2 + lam) * torch.sin(np.pi * X[:, 0]) * torch.sin(np.pi * X[:, 1])
          * torch.sin(np.pi * X[:, 2]))
    A = assembly.Assembly(ids, nm, dirichlet=dirichlet)
    rhs = A.assemble(ipr.apply(fx.contiguous()))
    return A, Hop, rhs, b


def test_gather_and_assemble_match_the_map():
    verts, ids = _mesh(3, True, seed=7, jitter=0.1)
    nm = 5
    l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm)
    A = assembly.Assembly(ids, nm, dirichlet=False)
    assert A.nglobal == ng
    rng = np.random.default_rng(2)
    g = rng.uniform(-1, 1, ng)
    loc = A.global_to_local(_dev(g)).cpu().numpy()
    assert np.array_equal(loc, sign * g[l2g])
    lv = rng.uniform(-1, 1, len(l2g))
    ref = np.zeros(ng)
    np.add.at(ref, l2g, sign * lv)
    got = A.assemble(_dev(lv)).cpu().numpy()
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("nm", [4, 6])
def test_helmsolve_matches_cpu_sparse_solve(nm):
    nq = nm + 1
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    print(f"[case] helmsolve rotated 2^3 mesh nm={nm}", flush=True)
    x_cpu, l2g, sign, u = cpu_solve(verts, ids, nm, nq, 10.0)
    A, Hop, rhs, _ = _gpu_problem(verts, ids, nm, nq, 10.0)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-13, max_iter=5000)
    xg = x.cpu().numpy()
    print(f"iterations {it}, residual {res:.2e}")
    assert res < 1e-13
    assert np.linalg.norm(xg - x_cpu) <= 1e-9 * np.linalg.norm(x_cpu)
    e_gpu = l2_error(verts, ids, nm, nq, xg, l2g, sign, u)
    e_cpu = l2_error(verts, ids, nm, nq, x_cpu, l2g, sign, u)
    assert abs(e_gpu - e_cpu) <= 1e-6 * e_cpu


def test_helmsolve_neumann_residual():
    """Natural boundary conditions (lambda > 0: SPD): the returned x satisfies
    Assemble(H(GlobalToLocal(x))) = b to the requested tolerance."""
    verts, ids = _mesh(3, True, seed=11, jitter=0.1)
    A, Hop, rhs, _ = _gpu_problem(verts, ids, 5, 6, 2.0, dirichlet=False)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-11)
    y = A.assemble(Hop.apply(A.global_to_local(x)))
    rel = (torch.linalg.vector_norm(y - rhs) / torch.linalg.vector_norm(rhs)).item()
    print(f"iterations {it}, residual {res:.2e}, check {rel:.2e}")
    assert rel < 1e-10


def test_helmsolve_reports_non_convergence():
    verts, ids = _mesh(2, False)
    A, Hop, rhs, _ = _gpu_problem(verts, ids, 5, 6, 1.0)
    with pytest.raises(H.MFError, match="maximum number of iterations"):
        A.helmsolve(Hop, rhs, tol=1e-14, max_iter=2)


# ---------------------------------------------------------------- quadrilaterals

def _gpu_quad_problem(verts, ids, nm, nq, lam, deformed=True, dirichlet=True):
    E = len(ids)
    b = H.basis_modified_a(nm, nq)
    geo = "Deformed" if deformed else "Regular"
    jac, df = H.geom_device("quad", _dev(verts), nq, b["z"], deformed)
    Hop = H.Operator(f"Helmholtz_Quad_{geo}", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator(f"IProduct_Quad_{geo}", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    _, f = manufactured(lam, 2)
    fx = f(O.coords_quad(verts, nq, b["z"]).T)  # host points of the quadrature grid
    A = assembly.Assembly(ids, nm, dirichlet=dirichlet, shape="quad")
    rhs = A.assemble(ipr.apply(_dev(fx)))
    return A, Hop, rhs


def test_quad_gather_and_assemble_match_the_map():
    verts, ids = _qmesh(4, True, seed=7, jitter=0.1)
    nm = 6
    l2g, sign, _, ng = assembly.assembly_map_quad(ids, nm)
    A = assembly.Assembly(ids, nm, dirichlet=False, shape="quad")
    assert A.nglobal == ng and A.nlocal == len(ids) * nm * nm
    rng = np.random.default_rng(3)
    g = rng.uniform(-1, 1, ng)
    assert np.array_equal(A.global_to_local(_dev(g)).cpu().numpy(), sign * g[l2g])
    lv = rng.uniform(-1, 1, len(l2g))
    ref = np.zeros(ng)
    np.add.at(ref, l2g, sign * lv)
    assert np.abs(A.assemble(_dev(lv)).cpu().numpy() - ref).max() < 1e-13


@pytest.mark.parametrize("nm,deformed", [(5, True), (8, True), (5, False)])
def test_quad_helmsolve_matches_cpu_sparse_solve(nm, deformed):
    nq = nm + 1
    if deformed:
        verts, ids = _qmesh(4, True, seed=3, jitter=0.1)
    else:  # the regular 32x32 quads of the measured quad workload
        verts, ids = _qmesh(32, False)
    print(f"[case] quad helmsolve nm={nm} deformed={deformed} E={len(ids)}", flush=True)
    x_cpu, l2g, sign, u = cpu_solve(verts, ids, nm, nq, 1.0, "quad")
    A, Hop, rhs = _gpu_quad_problem(verts, ids, nm, nq, 1.0, deformed)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-13, max_iter=5000)
    xg = x.cpu().numpy()
    print(f"iterations {it}, residual {res:.2e}")
    assert res < 1e-13
    assert np.linalg.norm(xg - x_cpu) <= 1e-9 * np.linalg.norm(x_cpu)
    e_gpu = l2_error(verts, ids, nm, nq, xg, l2g, sign, u, "quad")
    e_cpu = l2_error(verts, ids, nm, nq, x_cpu, l2g, sign, u, "quad")
    # relative L2 errors; at nm = 8 they are ~3e-10, so allow round-off in absolute terms
    assert abs(e_gpu - e_cpu) <= 1e-6 * e_cpu + 1e-12


def test_quad_helmsolve_rejects_a_hex_operator():
    verts, ids = _qmesh(2, False)
    A = assembly.Assembly(ids, 4, shape="quad")
    b = H.basis_modified_a(4, 5)
    Hop = H.Operator("Helmholtz_Hex_Regular", 4, 5, len(ids), basis=b)
    rhs = torch.zeros(A.nglobal, dtype=torch.float64, device="cuda")
    with pytest.raises(H.MFError, match="Helmholtz operator"):
        A.helmsolve(Hop, rhs)


def test_smoothing_a_nektar_field():
    """The reference's FieldConvert Helmholtz smoother (ProcessC0Projection.cpp:196-224:
    lambda = (2 pi / L)^2 with L = 0.6, forcing -lambda u, natural boundary conditions) on
    the mesh and field of its test Hex_channel_C0helmsmoothing, through the GPU BwdTrans,
    IProduct, assembly and CG: the solution equals a dense-element sparse direct solve, and
    the seven printed L2 / "L inf" lines of Hex_channel_C0helmsmoothing.tst are reproduced
    at their tolerances.  As in the reference, the expansion comes from the field file
    (hexahedra: nm GLL points per direction, MeshGraph.cpp:1179-1183), the norms are taken
    after re-expanding on nm evenly spaced points per direction (OutputFileBase.cpp:83-86)
    and "L inf" is the largest signed value (ExpList.cpp:3034-3037)."""
    from test_nektar_io import SMOOTHING_GOLDEN, equispaced_rule, printed
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    nmesh = N.read_mesh(os.path.join(gold, "Hex_channel_C0helmsmoothing.xml"))
    blocks = N.read_fld(os.path.join(gold, "Hex_channel_C0helmsmoothing.fld"))
    lam = (2 * np.pi / 0.6) ** 2
    got = {}
    for f in ("u", "v", "w", "p"):
        print(f"[case] C0 Helmholtz smoothing of field {f}", flush=True)
        X, ids, c, nm, nq = N.expansion_arrays(nmesh, blocks, f, points="field")
        E, M = len(ids), nm ** 3
        # CPU: dense element matrices assembled with the same map, every mode free
        l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm)
        rows, cols, vals, rhs = [], [], [], np.zeros(ng)
        for e in range(E):
            m = DR.element_mats("hex", nm, nq, X[e])
            g, s_ = l2g[e * M:(e + 1) * M], sign[e * M:(e + 1) * M].astype(float)
            rows.append(np.repeat(g, M))
            cols.append(np.tile(g, M))
            vals.append((DR.helmholtz_matrix(m, lam) * np.outer(s_, s_)).ravel())
            phys = m["B"] @ c[e * M:(e + 1) * M]
            np.add.at(rhs, g, s_ * (m["B"].T @ (m["W"] * m["J"] * lam * phys)))
        Ag = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                           shape=(ng, ng))
        x_cpu = spla.spsolve(Ag.tocsc(), rhs)
        # GPU
        b = H.basis_modified_a(nm, nq)
        vd = _dev(X)
        jac, df = H.geom_device("hex", vd, nq, b["z"], True)
        Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
        Hop.set_jac(jac)
        Hop.set_df(df)
        Hop.set_lambda(lam)
        ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E, basis=b)
        ipr.set_jac(jac)
        bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E, basis=b)
        A = assembly.Assembly(ids, nm, dirichlet=False)
        phys = bwd.apply(_dev(c))
        x, it, res = A.helmsolve(Hop, A.assemble(ipr.apply(lam * phys)), tol=1e-13)
        print(f"iterations {it}, residual {res:.2e}")
        assert np.linalg.norm(x.cpu().numpy() - x_cpu) <= 1e-10 * np.linalg.norm(x_cpu)
        # the output expansion: nm evenly spaced points per direction, Newton-Cotes weights
        ze, we = equispaced_rule(nm)
        B1, _ = DR.modified_a(nm, ze)
        beq = dict(bdata=np.ascontiguousarray(B1.T.ravel()), dbdata=np.zeros(nm * nm),
                   D=np.zeros(nm * nm), w=we, nm=nm, nq=nm)
        out = H.Operator("BwdTrans_Hex_Regular", nm, nm, E, basis=beq)
        ue = out.apply(A.global_to_local(x))
        jeq, _ = H.geom_device("hex", vd, nm, ze, True)
        W = torch.from_numpy(np.kron(np.kron(we, we), we)).cuda()
        wj = (jeq.view(E, -1) * W).reshape(-1)
        got[f] = (torch.sqrt(torch.dot(wj * ue, ue)).item(), ue.max().item())
        if f == "u":
            Xe = mesh.hex_quad_points(vd, ze)
            for a, k in zip("xyz", range(3)):
                col = Xe[..., k].reshape(-1)
                got[a] = (torch.sqrt(torch.dot(wj * col, col)).item(), col.max().item())
    for name, l2, tl2, linf, tlinf in SMOOTHING_GOLDEN:
        print(f"{name}: L2 {got[name][0]:.7g} (golden {l2}), L inf {got[name][1]:.7g} "
              f"(golden {linf})")
        assert abs(printed(got[name][0]) - l2) <= tl2, name
        assert abs(printed(got[name][1]) - linf) <= tlinf, name


def test_helmholtz3d_hex_golden_on_the_gpu():
    """The reference's Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex.tst (L2 error
    1.11818e-07, Linf 5.66374e-07, tolerance 1e-12) through the GPU path: geometry,
    IProduct of the forcing, assembly, the inhomogeneous Dirichlet lift and CG on the GPU,
    errors at the GLL points from the GPU BwdTrans."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    msh = N.read_mesh(os.path.join(gold, "Helmholtz3D_Hex_P6.xml"))
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm = msh.expansions[0]["nummodes"]
    nq, E, lam = nm + 1, len(ids), 1.0
    uex = lambda P: np.sin(np.pi / 2 * P[:, 0]) * np.sin(np.pi / 2 * P[:, 1]) * np.sin(np.pi / 2 * P[:, 2])
    b = H.basis_modified_a(nm, nq)
    jac, df = H.geom_device("hex", _dev(X), nq, b["z"], True)
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E, basis=b)
    P = mesh.hex_quad_points(_dev(X), b["z"]).reshape(-1, 3).cpu().numpy()
    A = assembly.Assembly(ids, nm, dirichlet=True)
    rhs = A.assemble(ipr.apply(_dev((lam + 3 * np.pi ** 2 / 4) * uex(P))))
    g = assembly.dirichlet_lift(X, ids, nm, uex)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-14, max_iter=5000, lift=g)
    d = bwd.apply(A.global_to_local(x)).cpu().numpy() - uex(P)
    wq = np.einsum("k,j,i->kji", b["w"], b["w"], b["w"]).ravel()
    wj = jac.cpu().numpy() * np.tile(wq, E)
    l2, linf = np.sqrt(np.sum(wj * d * d)), np.abs(d).max()
    print(f"iterations {it}, residual {res:.1e}, L2 {l2:.6e}, Linf {linf:.6e}")
    assert abs(l2 - 1.11818e-07) <= 1e-12
    assert abs(linf - 5.66374e-07) <= 1e-12


def test_helmholtz3d_hex_allbcs_golden_on_the_gpu():
    """Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_AllBCs.tst (L2 0.000416575, Linf
    0.000871589, tolerance 1e-12 on the printed values) through the GPU path: Dirichlet on
    two regions (the mask of the Dirichlet faces' closure and their lift), Neumann on two
    and Robin on two (the face integrals on the right-hand side and the Robin trace matrix
    inside the GPU conjugate gradients)."""
    from test_assembly_cpu import _allbcs_problem, _printed
    msh, uex, bc = _allbcs_problem()
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm = msh.expansions[0]["nummodes"]
    nq, E, lam = nm + 1, len(ids), 1.0
    print(f"[case] allbcs E={E} nm={nm}", flush=True)
    b = H.basis_modified_a(nm, nq)
    jac, df = H.geom_device("hex", _dev(X), nq, b["z"], True)
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E, basis=b)
    P = mesh.hex_quad_points(_dev(X), b["z"]).reshape(-1, 3).cpu().numpy()
    A = assembly.Assembly(ids, nm, dirichlet=True)
    g, dmask = assembly.dirichlet_lift_faces(X, ids, nm, uex, bc["dirichlet"])
    brhs, rr, rc, rv = assembly.boundary_terms(X, ids, nm, bc["neumann"], bc["robin"])
    A.set_dirichlet_mask(dmask)
    A.set_robin(rr, rc, rv)
    rhs = A.assemble(ipr.apply(_dev((lam + 3 * np.pi ** 2 / 4) * uex(P)))) + _dev(brhs)
    x, it, res = A.helmsolve(Hop, rhs, tol=1e-14, max_iter=5000, lift=g)
    d = bwd.apply(A.global_to_local(x)).cpu().numpy() - uex(P)
    wq = np.einsum("k,j,i->kji", b["w"], b["w"], b["w"]).ravel()
    wj = jac.cpu().numpy() * np.tile(wq, E)
    l2, linf = np.sqrt(np.sum(wj * d * d)), np.abs(d).max()
    print(f"iterations {it}, residual {res:.1e}, L2 {l2:.9e}, Linf {linf:.9e}")
    assert abs(_printed(l2) - 0.000416575) <= 1e-12
    assert abs(_printed(linf) - 0.000871589) <= 1e-12
    # removing the Robin matrix and the mask again restores the all-Dirichlet assembly
    A.set_robin([], [], [])
    A.set_dirichlet_mask(None)
    with pytest.raises(H.MFError):
        A.set_robin([0], [A.nglobal], [1.0])


# ---------------------------------------------------------------- variable order

def _runs(nmv):
    """(first, end, order) of each run of equal-order elements, in element order."""
    runs, e0 = [], 0
    for e in range(1, len(nmv) + 1):
        if e == len(nmv) or nmv[e] != nmv[e0]:
            runs.append((e0, e, int(nmv[e0])))
            e0 = e
    return runs


def _varp_gpu_problem(verts, ids, nmv, lam, shape):
    dim = 3 if shape == "hex" else 2
    S = "Hex" if shape == "hex" else "Quad"
    _, f = manufactured(lam, dim)
    ops, loc = [], []
    for e0, e1, nm in _runs(nmv):
        nq = nm + 1
        b = H.basis_modified_a(nm, nq)
        v = verts[e0:e1]
        jac, df = H.geom_device(shape, _dev(v), nq, b["z"], True)
        Hop = H.Operator(f"Helmholtz_{S}_Deformed", nm, nq, e1 - e0, basis=b)
        Hop.set_jac(jac)
        Hop.set_df(df)
        Hop.set_lambda(lam)
        ipr = H.Operator(f"IProduct_{S}_Deformed", nm, nq, e1 - e0, basis=b)
        ipr.set_jac(jac)
        pts = (O.coords_hex if shape == "hex" else O.coords_quad)(v, nq, b["z"]).T
        loc.append(ipr.apply(_dev(f(pts))))
        ops.append(Hop)
    A = assembly.Assembly(ids, nmv, dirichlet=True, shape=shape)
    return A, ops, A.assemble(torch.cat(loc).contiguous())


@pytest.mark.parametrize("shape", ["hex", "quad"])
def test_variable_order_gather_and_assemble(shape):
    from test_assembly_varp import _offsets  # noqa: F401  (same layout as the CPU tests)
    verts, ids = (_mesh if shape == "hex" else _qmesh)(3, True, seed=8, jitter=0.1)
    nmv = np.random.default_rng(9).integers(3, 7, len(ids))
    l2g, sign, _, ng = assembly.assembly_map(ids, nmv, shape)
    assert (sign == 0).any()
    A = assembly.Assembly(ids, nmv, dirichlet=False, shape=shape)
    assert A.nglobal == ng and A.nlocal == len(l2g)
    rng = np.random.default_rng(4)
    g = rng.uniform(-1, 1, ng)
    assert np.array_equal(A.global_to_local(_dev(g)).cpu().numpy(), sign * g[l2g])
    lv = rng.uniform(-1, 1, len(l2g))
    ref = np.zeros(ng)
    np.add.at(ref, l2g, sign * lv)
    assert np.abs(A.assemble(_dev(lv)).cpu().numpy() - ref).max() < 1e-13


@pytest.mark.parametrize("shape", ["hex", "quad"])
def test_variable_order_helmsolve_matches_cpu_sparse_solve(shape):
    from test_assembly_varp import varp_solve, varp_error
    if shape == "hex":
        verts, ids = _mesh(3, True, seed=3, jitter=0.05)
        nmv = np.repeat([4, 6, 5], 9)  # one order per z layer (elements are z-major)
    else:
        verts, ids = _qmesh(6, True, seed=3, jitter=0.05)
        nmv = np.repeat([3, 7, 5, 5, 8, 4], 6)  # one order per row
    print(f"[case] variable-order helmsolve {shape} orders {sorted(set(nmv.tolist()))}", flush=True)
    x_cpu, l2g, sign, u = varp_solve(verts, ids, nmv, 10.0, shape)
    A, ops, rhs = _varp_gpu_problem(verts, ids, nmv, 10.0, shape)
    x, it, res = A.helmsolve(ops, rhs, tol=1e-13, max_iter=5000)
    xg = x.cpu().numpy()
    print(f"operators {len(ops)}, iterations {it}, residual {res:.2e}")
    assert res < 1e-13
    assert np.linalg.norm(xg - x_cpu) <= 1e-9 * np.linalg.norm(x_cpu)
    e_gpu = varp_error(verts, ids, nmv, xg, l2g, sign, u, shape)
    e_cpu = varp_error(verts, ids, nmv, x_cpu, l2g, sign, u, shape)
    assert abs(e_gpu - e_cpu) <= 1e-6 * e_cpu


def test_variable_order_rejects_mismatched_operators():
    verts, ids = _mesh(2, False)
    nmv = np.repeat([4, 5], 4)
    A, ops, rhs = _varp_gpu_problem(verts, ids, nmv, 1.0, "hex")
    with pytest.raises(H.MFError, match="modes"):
        A.helmsolve(ops[::-1], rhs)
    with pytest.raises(H.MFError, match="cover"):
        A.helmsolve(ops[:1], rhs)


def test_preconditioner_follows_new_geometry():
    """Re-setting an operator's geometry rebuilds the cached assembled diagonal: the solve
    with the reused operator equals, bit for bit and in iterations, the solve with a fresh
    operator on the new geometry."""
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    nm, nq = 5, 6
    A, Hop, rhs, b = _gpu_problem(verts, ids, nm, nq, 10.0)
    A.helmsolve(Hop, rhs, tol=1e-12)  # caches the diagonal of the first geometry
    verts2 = np.ascontiguousarray(verts * np.array([1.0, 2.0, 0.5]))
    jac2, df2 = H.geom_device("hex", _dev(verts2), nq, b["z"], True)
    Hop.set_jac(jac2)
    Hop.set_df(df2)
    x1, it1, _ = A.helmsolve(Hop, rhs, tol=1e-12)
    A2, Hop2, _, _ = _gpu_problem(verts2, ids, nm, nq, 10.0)
    x2, it2, _ = A2.helmsolve(Hop2, rhs, tol=1e-12)
    print(f"iterations {it1} / {it2}")
    assert it1 == it2
    assert torch.equal(x1, x2)


def test_preconditioner_follows_a_recreated_operator():
    """An operator destroyed and created again (its handle may land at the freed address)
    on new geometry with the same lambda: the cached assembled diagonal is keyed on a
    process-wide geometry id, so the reused assembly solves exactly as a fresh one."""
    import gc
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    nm, nq = 5, 6
    A, Hop, rhs, b = _gpu_problem(verts, ids, nm, nq, 10.0)
    A.helmsolve(Hop, rhs, tol=1e-12)  # caches the diagonal of the first geometry
    for s in (0.5, 2.0):
        del Hop
        gc.collect()
        verts2 = np.ascontiguousarray(verts * np.array([1.0, s, 1.0]))
        jac2, df2 = H.geom_device("hex", _dev(verts2), nq, b["z"], True)
        Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, len(ids), basis=b)
        Hop.set_jac(jac2)
        Hop.set_df(df2)
        Hop.set_lambda(10.0)
        x1, it1, _ = A.helmsolve(Hop, rhs, tol=1e-12)
        A2, Hop2, _, _ = _gpu_problem(verts2, ids, nm, nq, 10.0)
        x2, it2, _ = A2.helmsolve(Hop2, rhs, tol=1e-12)
        print(f"scale {s}: iterations {it1} / {it2}")
        assert it1 == it2
        assert torch.equal(x1, x2)
