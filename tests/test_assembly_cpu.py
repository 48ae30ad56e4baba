"""C0 global numbering (host part of the library) checked on the CPU.

* counts: a box of n^3 hexes with nm modes has (n(nm-1)+1)^3 global modes, for any local
  orientation of the elements;
* continuity: a random global vector, scattered with the map and evaluated with the oracle
  BwdTrans, takes one value at every physical quadrature point shared by several elements,
  with every element listed in a random rotation of its reference frame (edge reversals,
  face flips and face transposes all occur);
* the discrete Helmholtz problem assembled with the map from independent dense element
  matrices (tests/dense_ref.py) converges spectrally to a manufactured solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import dense_ref as DR
import oracle as O
from ithaca_hip import assembly, mesh


def _mesh(n, rotate, seed=0, jitter=0.0):
    rng = np.random.default_rng(seed)
    verts = mesh.hex_box(n, jitter, rng)
    ids = mesh.hex_box_ids(n)
    if rotate:
        verts, ids = mesh.rotate_elements(verts, ids, rng)
    return verts, ids


@pytest.mark.parametrize("n,nm,rotate", [(2, 4, False), (3, 5, True), (2, 8, True)])
def test_global_count(n, nm, rotate):
    verts, ids = _mesh(n, rotate)
    l2g, sign, bnd, ng = assembly.assembly_map_hex(ids, nm)
    assert ng == (n * (nm - 1) + 1) ** 3
    assert set(np.unique(l2g)) == set(range(ng))
    if not rotate:
        assert np.all(sign == 1)
    # boundary modes: everything except the interior (n(nm-1)-1)^3 block
    assert len(np.unique(l2g[bnd == 1])) == ng - (n * (nm - 1) - 1) ** 3


@pytest.mark.parametrize("nm,nq", [(4, 5), (6, 7)])
def test_continuity_rotated_elements(nm, nq):
    verts, ids = _mesh(3, True, seed=nm, jitter=0.1)
    E = len(ids)
    l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm)
    assert (sign < 0).any()
    x = np.random.default_rng(1).uniform(-1, 1, ng)
    loc = sign * x[l2g]
    b = O.basis(nm, nq)
    vals = O.bwdtrans("hex", b, E, loc)
    xyz = O.coords_hex(verts, nq, b["z"]).T  # (E*Q, 3)
    key = np.round(xyz * 1e9).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.ravel()
    lo = np.full(inv.max() + 1, np.inf)
    hi = np.full(inv.max() + 1, -np.inf)
    np.minimum.at(lo, inv, vals)
    np.maximum.at(hi, inv, vals)
    shared = np.bincount(inv) > 1
    assert shared.sum() > 0
    assert np.max(hi[shared] - lo[shared]) < 1e-12


def helmholtz_system(verts, ids, nm, nq, lam, f_exact, shape="hex"):
    """Global CSR matrix and rhs (Dirichlet modes removed) from dense element matrices."""
    E = len(ids)
    M = nm ** (3 if shape == "hex" else 2)
    l2g, sign, bnd, ng = assembly.assembly_map(ids, nm, shape)
    rows, cols, vals = [], [], []
    rhs = np.zeros(ng)
    for e in range(E):
        m = DR.element_mats(shape, nm, nq, verts[e])
        He = DR.helmholtz_matrix(m, lam)
        g = l2g[e * M:(e + 1) * M]
        s = sign[e * M:(e + 1) * M].astype(float)
        Hs = He * np.outer(s, s)
        rows.append(np.repeat(g, M))
        cols.append(np.tile(g, M))
        vals.append(Hs.ravel())
        fe = f_exact(m["X"])
        np.add.at(rhs, g, s * (m["B"].T @ (m["W"] * m["J"] * fe)))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ng, ng))
    dir_mask = np.zeros(ng, dtype=bool)
    dir_mask[np.unique(l2g[bnd == 1])] = True
    return A, rhs, dir_mask, l2g, sign


def manufactured(lam, dim=3):
    """u = prod_c sin(pi x_c) (zero on the boundary of the unit box), f = -lap u + lam u."""
    u = lambda X: np.prod(np.sin(np.pi * X[:, :dim]), axis=1)
    f = lambda X: (dim * np.pi ** 2 + lam) * u(X)
    return u, f


def cpu_solve(verts, ids, nm, nq, lam, shape="hex"):
    u, f = manufactured(lam, 3 if shape == "hex" else 2)
    A, rhs, dm, l2g, sign = helmholtz_system(verts, ids, nm, nq, lam, f, shape)
    free = ~dm
    x = np.zeros(len(rhs))
    x[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free])
    return x, l2g, sign, u


def l2_error(verts, ids, nm, nq, x, l2g, sign, u, shape="hex"):
    M = nm ** (3 if shape == "hex" else 2)
    num = den = 0.0
    for e in range(len(ids)):
        m = DR.element_mats(shape, nm, nq, verts[e])
        ue = m["B"] @ (sign[e * M:(e + 1) * M] * x[l2g[e * M:(e + 1) * M]])
        ex = u(m["X"])
        num += np.sum(m["W"] * m["J"] * (ue - ex) ** 2)
        den += np.sum(m["W"] * m["J"] * ex ** 2)
    return np.sqrt(num / den)


def test_manufactured_solution_converges():
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    errs = []
    for nm in (4, 5, 6, 7):
        x, l2g, sign, u = cpu_solve(verts, ids, nm, nm + 1, 10.0)
        errs.append(l2_error(verts, ids, nm, nm + 1, x, l2g, sign, u))
    print("L2 errors", errs)
    assert errs[-1] < 1e-4
    assert all(errs[i + 1] < 0.5 * errs[i] for i in range(len(errs) - 1))


# ---------------------------------------------------------------- quadrilaterals

def _qmesh(n, rotate, seed=0, jitter=0.0):
    rng = np.random.default_rng(seed)
    verts = mesh.quad_box(n, jitter, rng)
    ids = mesh.quad_box_ids(n)
    if rotate:
        verts, ids = mesh.rotate_quads(verts, ids, rng)
    return verts, ids


@pytest.mark.parametrize("n,nm,rotate", [(3, 4, False), (4, 5, True), (2, 8, True)])
def test_quad_global_count(n, nm, rotate):
    verts, ids = _qmesh(n, rotate)
    l2g, sign, bnd, ng = assembly.assembly_map_quad(ids, nm)
    assert ng == (n * (nm - 1) + 1) ** 2
    assert set(np.unique(l2g)) == set(range(ng))
    if not rotate:
        assert np.all(sign == 1)
    assert len(np.unique(l2g[bnd == 1])) == ng - (n * (nm - 1) - 1) ** 2


@pytest.mark.parametrize("nm,nq", [(4, 5), (7, 8)])
def test_quad_continuity_rotated_elements(nm, nq):
    verts, ids = _qmesh(4, True, seed=nm, jitter=0.15)
    E = len(ids)
    l2g, sign, _, ng = assembly.assembly_map_quad(ids, nm)
    assert (sign < 0).any()
    x = np.random.default_rng(1).uniform(-1, 1, ng)
    vals = O.bwdtrans("quad", O.basis(nm, nq), E, sign * x[l2g])
    xy = O.coords_quad(verts, nq, O.basis(nm, nq)["z"]).T
    _, inv = np.unique(np.round(xy * 1e9).astype(np.int64), axis=0, return_inverse=True)
    inv = inv.ravel()
    lo = np.full(inv.max() + 1, np.inf)
    hi = np.full(inv.max() + 1, -np.inf)
    np.minimum.at(lo, inv, vals)
    np.maximum.at(hi, inv, vals)
    shared = np.bincount(inv) > 1
    assert shared.sum() > 0
    assert np.max(hi[shared] - lo[shared]) < 1e-12


def test_quad_assembly_rejects_bad_meshes():
    ids = mesh.quad_box_ids(2)
    bad = ids.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(Exception, match="degenerate edge"):
        assembly.assembly_map_quad(bad, 4)
    triple = np.concatenate([ids, ids[:1], ids[:1]])
    with pytest.raises(Exception, match="more than two elements"):
        assembly.assembly_map_quad(triple, 4)


def test_quad_manufactured_solution_converges():
    verts, ids = _qmesh(3, True, seed=5, jitter=0.1)
    errs = []
    for nm in (4, 5, 6, 7, 8):
        x, l2g, sign, u = cpu_solve(verts, ids, nm, nm + 1, 1.0, "quad")
        errs.append(l2_error(verts, ids, nm, nm + 1, x, l2g, sign, u, "quad"))
    print("L2 errors", errs)
    assert errs[-1] < 1e-6
    assert all(errs[i + 1] < 0.5 * errs[i] for i in range(len(errs) - 1))


# ------------------------------------------------ the reference's solver-level golden

def test_helmholtz3d_hex_golden():
    """Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex.tst: the session
    Helmholtz3D_Hex_P6.xml (8 hexes on the unit cube, NUMMODES 6, lambda = 1, Dirichlet
    u = sin(pi x/2) sin(pi y/2) sin(pi z/2) on every boundary face, forcing
    -(lambda + 3 pi^2/4) u) solved by the reference's Helmholtz3D gives L2 error 1.11818e-07
    and Linf error 5.66374e-07 (tolerance 1e-12).  Here: the mesh read with nektar_io, the
    C0 map, the Dirichlet values by the boundary-constrained projection of
    assembly.dirichlet_lift, dense element matrices and a sparse direct solve."""
    import os
    from ithaca_hip import nektar_io
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    mesh = nektar_io.read_mesh(os.path.join(gold, "Helmholtz3D_Hex_P6.xml"))
    X, ids = mesh.element_coords(), mesh.vertex_ids()
    nm = mesh.expansions[0]["nummodes"]
    nq, M, lam = nm + 1, nm ** 3, 1.0
    uex = lambda P: np.sin(np.pi / 2 * P[:, 0]) * np.sin(np.pi / 2 * P[:, 1]) * np.sin(np.pi / 2 * P[:, 2])
    g = assembly.dirichlet_lift(X, ids, nm, uex)
    l2g, sign, bnd, ng = assembly.assembly_map_hex(ids, nm)
    dmask = np.zeros(ng, dtype=bool)
    dmask[np.unique(l2g[bnd == 1])] = True
    assert np.all(g[~dmask] == 0)
    mats = [DR.element_mats("hex", nm, nq, X[e]) for e in range(len(ids))]
    rows, cols, vals, rhs = [], [], [], np.zeros(ng)
    for e, m in enumerate(mats):
        gi, s = l2g[e * M:(e + 1) * M], sign[e * M:(e + 1) * M].astype(float)
        rows.append(np.repeat(gi, M))
        cols.append(np.tile(gi, M))
        vals.append((DR.helmholtz_matrix(m, lam) * np.outer(s, s)).ravel())
        np.add.at(rhs, gi, s * (m["B"].T @ (m["W"] * m["J"] * (lam + 3 * np.pi ** 2 / 4) * uex(m["X"]))))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ng, ng))
    x = g.copy()
    free = ~dmask
    x[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free] - A[free][:, dmask] @ g[dmask])
    l2 = linf = 0.0
    for e, m in enumerate(mats):
        d = m["B"] @ (sign[e * M:(e + 1) * M] * x[l2g[e * M:(e + 1) * M]]) - uex(m["X"])
        l2 += np.sum(m["W"] * m["J"] * d * d)
        linf = max(linf, np.abs(d).max())
    print(f"L2 {np.sqrt(l2):.6e}  Linf {linf:.6e}")
    assert abs(np.sqrt(l2) - 1.11818e-07) <= 1e-12
    assert abs(linf - 5.66374e-07) <= 1e-12


# ---------------------------------------------------------------- per-direction orders

@pytest.mark.parametrize("n,nm3", [(2, (4, 6, 8)), (3, (3, 5, 2))])
def test_global_count_per_direction_orders(n, nm3):
    """Aligned elements with nm3[d] modes along direction d: the global space is the tensor
    product of 1D spaces with n (nm_d - 1) + 1 modes."""
    verts, ids = _mesh(n, False)
    l2g, sign, bnd, ng = assembly.assembly_map_hex(ids, nm3)
    assert ng == np.prod([n * (m - 1) + 1 for m in nm3])
    assert len(l2g) == n ** 3 * np.prod(nm3)
    assert set(np.unique(l2g)) == set(range(ng))
    assert np.all(sign == 1)
    interior = np.prod([n * (m - 1) - 1 for m in nm3])
    assert len(np.unique(l2g[bnd == 1])) == ng - interior
    # with equal orders the per-direction map is the isotropic one bit for bit
    if len(set(nm3)) == 1:
        ref = assembly.assembly_map_hex(ids, nm3[0])
        assert np.array_equal(l2g, ref[0]) and np.array_equal(sign, ref[1])


def test_continuity_per_direction_orders_rotated_elements():
    """Every element with (4, 6, 8) modes along its local directions, elements in random
    rotations of their frames: shared faces and edges get the fewest modes of their
    elements along each face axis (the rest sign 0), and a random global field, evaluated
    from each element with dense per-direction tables on 9 points per direction, takes one
    value at every shared point."""
    verts, ids = _mesh(3, True, seed=4, jitter=0.1)
    nm3, nq3 = (4, 6, 8), (9, 9, 9)
    E, M = len(ids), int(np.prod(nm3))
    l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm3)
    assert (sign < 0).any() and (sign == 0).any()
    x = np.random.default_rng(1).uniform(-1, 1, ng)
    loc = sign * x[l2g]
    vals, pts = [], []
    for e in range(E):
        m = DR.element_mats("hex", nm3, nq3, verts[e])
        vals.append(m["B"] @ loc[e * M:(e + 1) * M])
        pts.append(m["X"])
    vals, pts = np.concatenate(vals), np.concatenate(pts)
    key = np.round(pts * 1e9).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.ravel()
    lo = np.full(inv.max() + 1, np.inf)
    hi = np.full(inv.max() + 1, -np.inf)
    np.minimum.at(lo, inv, vals)
    np.maximum.at(hi, inv, vals)
    shared = np.bincount(inv) > 1
    assert shared.sum() > 0
    assert np.max(hi[shared] - lo[shared]) < 1e-12


def test_helmholtz3d_hex_heterogeneous_golden():
    """Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_Heterogeneous.tst: the session
    Helmholtz3D_Hex_Heterogeneous.xml (8 hexes on the unit cube, NUMMODES 4,6,8 and
    NUMPOINTS 5,7,9 GLL along the local directions, lambda = 1, the Dirichlet data and
    forcing of Helmholtz3D_CG_Hex) gives L2 error 5.23385e-05 and Linf error 0.00014837
    (tolerance 1e-12).  Here: the per-direction map, the boundary-constrained projection of
    the Dirichlet data with each face's two tables, dense per-direction element matrices
    and a sparse direct solve."""
    import os
    from ithaca_hip import nektar_io
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    msh = nektar_io.read_mesh(os.path.join(gold, "Helmholtz3D_Hex_Heterogeneous.xml"))
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm3, nq3 = msh.expansions[0]["nummodes"], msh.expansions[0]["numpoints"]
    assert (nm3, nq3) == ((4, 6, 8), (5, 7, 9))
    M, lam = int(np.prod(nm3)), 1.0
    uex = lambda P: np.sin(np.pi / 2 * P[:, 0]) * np.sin(np.pi / 2 * P[:, 1]) * np.sin(np.pi / 2 * P[:, 2])
    g = assembly.dirichlet_lift(X, ids, nm3, uex, nq=nq3)
    l2g, sign, bnd, ng = assembly.assembly_map_hex(ids, nm3)
    dmask = np.zeros(ng, dtype=bool)
    dmask[np.unique(l2g[bnd == 1])] = True
    assert np.all(g[~dmask] == 0)
    mats = [DR.element_mats("hex", nm3, nq3, X[e]) for e in range(len(ids))]
    rows, cols, vals, rhs = [], [], [], np.zeros(ng)
    for e, m in enumerate(mats):
        gi, s = l2g[e * M:(e + 1) * M], sign[e * M:(e + 1) * M].astype(float)
        rows.append(np.repeat(gi, M))
        cols.append(np.tile(gi, M))
        vals.append((DR.helmholtz_matrix(m, lam) * np.outer(s, s)).ravel())
        np.add.at(rhs, gi, s * (m["B"].T @ (m["W"] * m["J"] * (lam + 3 * np.pi ** 2 / 4) * uex(m["X"]))))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ng, ng))
    x = g.copy()
    free = ~dmask
    x[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free] - A[free][:, dmask] @ g[dmask])
    l2 = linf = 0.0
    for e, m in enumerate(mats):
        d = m["B"] @ (sign[e * M:(e + 1) * M] * x[l2g[e * M:(e + 1) * M]]) - uex(m["X"])
        l2 += np.sum(m["W"] * m["J"] * d * d)
        linf = max(linf, np.abs(d).max())
    print(f"L2 {np.sqrt(l2):.6e}  Linf {linf:.6e}")
    assert abs(float(f"{np.sqrt(l2):.6g}") - 5.23385e-05) <= 1e-12
    assert abs(float(f"{linf:.6g}") - 0.00014837) <= 1e-12


# ---------------------------------------------------------------- mixed boundary conditions

def _allbcs_problem():
    """Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_AllBCs.tst: the session
    Helmholtz3D_Hex_AllBCs_P6.xml (27 hexes, NUMMODES 4, lambda = 1, u = sin(pi x/2)
    sin(pi y/2) sin(pi z/2)) with Dirichlet regions C[1], C[5], Neumann C[2], C[6] and Robin
    C[3] (du/dn + 2u = g) and C[4] (du/dn + 3u = g); the condition values as the session
    states them."""
    import os
    from ithaca_hip import nektar_io
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")
    msh = nektar_io.read_mesh(os.path.join(gold, "Helmholtz3D_Hex_AllBCs_P6.xml"))
    pi = np.pi
    S = lambda P, k: np.sin(pi / 2 * P[:, k])
    C = lambda P, k: np.cos(pi / 2 * P[:, k])
    uex = lambda P: S(P, 0) * S(P, 1) * S(P, 2)
    F = {c: msh.boundary_faces(c) for c in range(1, 7)}
    bc = dict(dirichlet=F[1] + F[5],
              neumann=[(F[2], lambda P: -S(P, 0) * C(P, 1) * S(P, 2) * pi / 2),
                       (F[6], lambda P: S(P, 0) * S(P, 1) * C(P, 2) * pi / 2)],
              robin=[(F[3], 2.0, lambda P: C(P, 0) * S(P, 1) * S(P, 2) * pi / 2 + 2 * uex(P)),
                     (F[4], 3.0, lambda P: S(P, 0) * C(P, 1) * S(P, 2) * pi / 2 + 3 * uex(P))])
    return msh, uex, bc


def _printed(v):
    """The value as the reference's tester reads it from the program's output (the default
    six significant digits of the stream)."""
    return float(f"{v:.6g}")


def test_helmholtz3d_hex_allbcs_golden():
    """Helmholtz3D_CG_Hex_AllBCs.tst: L2 error 0.000416575, Linf 0.000871589 (tolerance
    1e-12 on the printed values).  Here: the Dirichlet projection on the Dirichlet faces
    only (dirichlet_lift_faces), the Neumann and Robin face integrals and the Robin trace
    matrix (boundary_terms), dense element matrices and a sparse direct solve."""
    msh, uex, bc = _allbcs_problem()
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm = msh.expansions[0]["nummodes"]
    nq, M, lam = nm + 1, nm ** 3, 1.0
    g, dmask = assembly.dirichlet_lift_faces(X, ids, nm, uex, bc["dirichlet"])
    brhs, rr, rc, rv = assembly.boundary_terms(X, ids, nm, bc["neumann"], bc["robin"])
    l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm)
    mats = [DR.element_mats("hex", nm, nq, X[e]) for e in range(len(ids))]
    rows, cols, vals, rhs = [rr], [rc], [rv], brhs.copy()
    for e, m in enumerate(mats):
        gi, s = l2g[e * M:(e + 1) * M], sign[e * M:(e + 1) * M].astype(float)
        rows.append(np.repeat(gi, M))
        cols.append(np.tile(gi, M))
        vals.append((DR.helmholtz_matrix(m, lam) * np.outer(s, s)).ravel())
        np.add.at(rhs, gi, s * (m["B"].T @ (m["W"] * m["J"] * (lam + 3 * np.pi ** 2 / 4) * uex(m["X"]))))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ng, ng))
    x = g.copy()
    free = ~dmask
    x[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free] - A[free][:, dmask] @ g[dmask])
    l2 = linf = 0.0
    for e, m in enumerate(mats):
        d = m["B"] @ (sign[e * M:(e + 1) * M] * x[l2g[e * M:(e + 1) * M]]) - uex(m["X"])
        l2 += np.sum(m["W"] * m["J"] * d * d)
        linf = max(linf, np.abs(d).max())
    print(f"L2 {np.sqrt(l2):.9e}  Linf {linf:.9e}")
    assert abs(_printed(np.sqrt(l2)) - 0.000416575) <= 1e-12
    assert abs(_printed(linf) - 0.000871589) <= 1e-12


def test_boundary_terms_integrate_the_faces():
    """With g = 1 the Neumann integrals against the constant field's coefficients give the
    areas of the faces (computed from the vertices as two triangles each), the Robin matrix
    is symmetric and 1^T R 1 = alpha * area."""
    msh, _, bc = _allbcs_problem()
    X, ids = msh.element_coords(), msh.vertex_ids()
    nm = 4
    one = lambda P: np.ones(len(P))
    g1, _ = assembly.dirichlet_lift_faces(X, ids, nm, one, [(e, f) for e in range(len(ids))
                                                            for f in range(6)])
    # the constant field: every vertex mode 1, every other mode 0
    l2g, sign, _, ng = assembly.assembly_map_hex(ids, nm)
    c1 = np.zeros(ng)
    loc = np.arange(nm ** 3).reshape(nm, nm, nm)
    vert = loc[:2, :2, :2].ravel()
    for e in range(len(ids)):
        c1[l2g[e * nm ** 3 + vert]] = 1.0
    assert np.allclose(g1[np.abs(g1) > 0], c1[np.abs(g1) > 0])
    corners = {0: (0, 1, 2, 3), 1: (0, 1, 5, 4), 2: (1, 2, 6, 5), 3: (3, 2, 6, 7),
               4: (0, 3, 7, 4), 5: (4, 5, 6, 7)}
    for faces in (bc["dirichlet"], bc["robin"][0][0]):
        area = 0.0
        for e, f in faces:
            a, b, c, d = (X[e][v] for v in corners[f])
            area += 0.5 * (np.linalg.norm(np.cross(b - a, c - a)) + np.linalg.norm(np.cross(c - a, d - a)))
        rhs, rr, rc, rv = assembly.boundary_terms(X, ids, nm, [(faces, one)], [(faces, 2.5, one)])
        assert abs(rhs @ c1 - 2 * area) < 1e-12 * area   # the Neumann and the Robin g = 1
        R = sp.csr_matrix((rv, (rr, rc)), shape=(ng, ng))
        assert abs(c1 @ (R @ c1) - 2.5 * area) < 1e-12 * area
        assert abs(R - R.T).max() < 1e-15
