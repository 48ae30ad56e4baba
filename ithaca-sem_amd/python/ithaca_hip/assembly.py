"""C0 global assembly and the matrix-free Helmholtz solve (include/ithaca_mf.h).

The reference assembles with AssemblyMapCG (MultiRegions/AssemblyMap/AssemblyMapCG.cpp:
GlobalToLocal :2853, Assemble :2885) and solves with NekLinSysIterCG
(LibUtilities/LinearAlgebra/NekLinSysIterCG.cpp:104) on
GlobalLinSysIterativeFull::v_DoMatrixMultiply (MultiRegions/GlobalLinSysIterativeFull.cpp:215).
"""
import ctypes

import numpy as np
import torch

from . import MFError, _stream_ptr, lib

_SET = False


def _lib():
    global _SET
    L = lib()
    if not _SET:
        vp, i, l_, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        for shape in ("hex", "quad"):
            getattr(L, f"imf_assembly_map_{shape}").argtypes = [i, i, vp, vp, vp, vp,
                                                                 ctypes.POINTER(l_)]
            getattr(L, f"imf_assembly_create_{shape}").argtypes = [i, i, vp, i,
                                                                    ctypes.POINTER(vp)]
        L.imf_assembly_map_varp.argtypes = [i, i, vp, vp, vp, vp, vp, ctypes.POINTER(l_)]
        L.imf_assembly_create_varp.argtypes = [i, i, vp, vp, i, ctypes.POINTER(vp)]
        L.imf_assembly_map_aniso.argtypes = [i, vp, vp, vp, vp, vp, ctypes.POINTER(l_)]
        L.imf_assembly_create_aniso.argtypes = [i, vp, vp, i, ctypes.POINTER(vp)]
        L.imf_helmsolve_cg_groups_device.argtypes = [vp, i, vp, vp, vp, d, i, ctypes.POINTER(i),
                                                     ctypes.POINTER(d), vp]
        L.imf_assembly_nglobal.argtypes = [vp]
        L.imf_assembly_nglobal.restype = l_
        L.imf_assembly_destroy.argtypes = [vp]
        L.imf_global_to_local_device.argtypes = [vp, vp, vp, vp]
        L.imf_assemble_device.argtypes = [vp, vp, vp, vp]
        L.imf_helmsolve_cg_device.argtypes = [vp, vp, vp, vp, d, i, ctypes.POINTER(i),
                                              ctypes.POINTER(d), vp]
        L.imf_assembly_last_error.restype = ctypes.c_char_p
        L.imf_assembly_set_dirichlet_mask.argtypes = [vp, vp]
        L.imf_assembly_set_robin.argtypes = [vp, l_, vp, vp, vp]
        _SET = True
    return L


def _check(st):
    if st != 0:
        msg = _lib().imf_assembly_last_error().decode() or lib().imf_last_error().decode()
        raise MFError(msg)


def _nverts(shape):
    if shape not in ("hex", "quad"):
        raise MFError(f"assembly: unknown shape {shape!r} (hex or quad)")
    return 8 if shape == "hex" else 4


def _orders(nm, E):
    """Per-element orders as int32 (None for a uniform int order)."""
    if np.ndim(nm) == 0:
        return None
    nmv = np.ascontiguousarray(nm, dtype=np.int32)
    if nmv.shape != (E,):
        raise MFError(f"assembly: {E} per-element orders expected, got shape {nmv.shape}")
    return nmv


def _orders3(nm, E):
    """Per-direction orders of hexahedra as an (E, 3) int32 array, from a tuple of three
    (every element; a list or array of length E gives per-element orders instead) or an
    (E, 3) array; None when nm is not per direction."""
    if isinstance(nm, tuple):
        if len(nm) != 3:
            raise MFError("assembly: per-direction orders are a tuple (nm0, nm1, nm2)")
        return np.ascontiguousarray(np.tile(np.asarray(nm, dtype=np.int32), (E, 1)))
    a = np.asarray(nm)
    if a.ndim == 2:
        if a.shape != (E, 3):
            raise MFError(f"assembly: per-direction orders must be (nElmt, 3), got {a.shape}")
        return np.ascontiguousarray(a, dtype=np.int32)
    return None


def assembly_map(vertex_ids, nm, shape):
    """(l2g, sign, on_boundary, nglobal) of the C0 numbering (host only, no device).  nm is
    the order of every element, or an array of per-element orders (variable order: modes
    beyond a shared edge's or face's order get l2g 0 and sign 0)."""
    ids = np.ascontiguousarray(vertex_ids, dtype=np.int64)
    if ids.ndim != 2 or ids.shape[1] != _nverts(shape):
        raise MFError(f"assembly: vertex ids must be (nElmt, {_nverts(shape)}) for {shape}")
    E = ids.shape[0]
    dim = 3 if shape == "hex" else 2
    nm3 = _orders3(nm, E) if shape == "hex" else None
    if nm3 is not None:
        n = int(np.sum(np.prod(nm3.astype(np.int64), axis=1)))
        l2g = np.empty(n, dtype=np.int32)
        sign = np.empty(n, dtype=np.int8)
        bnd = np.empty(n, dtype=np.uint8)
        ng = ctypes.c_long()
        _check(_lib().imf_assembly_map_aniso(E, nm3.ctypes.data_as(ctypes.c_void_p),
                                             ids.ctypes.data_as(ctypes.c_void_p),
                                             l2g.ctypes.data_as(ctypes.c_void_p),
                                             sign.ctypes.data_as(ctypes.c_void_p),
                                             bnd.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ng)))
        return l2g, sign, bnd, ng.value
    nmv = _orders(nm, E)
    n = int(np.sum(nmv.astype(np.int64) ** dim)) if nmv is not None else E * nm ** dim
    l2g = np.empty(n, dtype=np.int32)
    sign = np.empty(n, dtype=np.int8)
    bnd = np.empty(n, dtype=np.uint8)
    ng = ctypes.c_long()
    ptrs = [ids.ctypes.data_as(ctypes.c_void_p), l2g.ctypes.data_as(ctypes.c_void_p),
            sign.ctypes.data_as(ctypes.c_void_p), bnd.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(ng)]
    if nmv is None:
        _check(getattr(_lib(), f"imf_assembly_map_{shape}")(E, nm, *ptrs))
    else:
        _check(_lib().imf_assembly_map_varp(dim, E, nmv.ctypes.data_as(ctypes.c_void_p), *ptrs))
    return l2g, sign, bnd, ng.value


def assembly_map_hex(vertex_ids, nm):
    return assembly_map(vertex_ids, nm, "hex")


def assembly_map_quad(vertex_ids, nm):
    return assembly_map(vertex_ids, nm, "quad")


# ---------------------------------------------------------------- Dirichlet data

def _tables(nm, nq):
    from . import basis_modified_a
    b = basis_modified_a(nm, nq)
    return np.asarray(b["z"]), np.asarray(b["w"]), np.asarray(b["bdata"]).reshape(nm, nq).T


def _constrained_1d(vals, B, w):
    """StdSegExp::v_FwdTrans_BndConstrained (StdRegions/StdSegExp.cpp:319-390): vertex
    modes = end values, interior modes by the interior block of the mass matrix with the
    vertex part removed."""
    nm = B.shape[1]
    c = np.zeros(nm)
    c[0], c[1] = vals[0], vals[-1]
    if nm > 2:
        M1 = B.T @ (w[:, None] * B)
        rhs = B.T @ (w * vals) - M1[:, :2] @ c[:2]
        c[2:] = np.linalg.solve(M1[2:, 2:], rhs[2:])
    return c


def _constrained_2d(F, B, w, Bs=None, ws=None):
    """StdQuadExp::v_FwdTrans_BndConstrained (StdRegions/StdQuadExp.cpp:260-354) on a face
    grid F[j, i] (i along the first tangential direction, tables B, w; j along the second,
    tables Bs, ws, the same as the first by default): vertices, then each edge by the 1D
    constrained projection, then the interior by the interior mass block.  Returns
    c[q, p].  The face Jacobian of a straight-sided (parallelogram) face is constant and
    cancels."""
    Bs = B if Bs is None else Bs
    ws = w if ws is None else ws
    nmi, nmj = B.shape[1], Bs.shape[1]
    c = np.zeros((nmj, nmi))
    c[0, :] = _constrained_1d(F[0, :], B, w)
    c[1, :] = _constrained_1d(F[-1, :], B, w)
    c[:, 0] = _constrained_1d(F[:, 0], Bs, ws)
    c[:, 1] = _constrained_1d(F[:, -1], Bs, ws)
    if nmi > 2 and nmj > 2:
        B2 = np.kron(Bs, B)
        W2 = np.kron(ws, w)
        M2 = B2.T @ (W2[:, None] * B2)
        it = np.zeros((nmj, nmi), dtype=bool)
        it[2:, 2:] = True
        it = it.ravel()
        cb = c.ravel().copy()
        cb[it] = 0.0
        rhs = B2.T @ (W2 * F.ravel()) - M2 @ cb
        cb[it] = np.linalg.solve(M2[np.ix_(it, it)], rhs[it])
        c = cb.reshape(nmj, nmi)
    return c


def _points(coords, z):
    """Physical GLL points of straight-sided elements: (E, nq, nq[, nq], dim), index
    [k][j][i] (hex) or [j][i] (quad), from the vertices in local order.  z: the points of
    every direction, or (hexes) a list of the three directions' points."""
    coords = np.asarray(coords, dtype=np.float64)
    nv, dim = coords.shape[1], coords.shape[2]
    if nv == 8:
        zs = z if isinstance(z, (list, tuple)) else [z, z, z]
        lo = [0.5 * (1.0 - np.asarray(x)) for x in zs]
        hi = [0.5 * (1.0 + np.asarray(x)) for x in zs]
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        P = 0.0
        for v, (a, b, c) in enumerate(corners):
            f = np.einsum("k,j,i->kji", hi[2] if c else lo[2], hi[1] if b else lo[1],
                          hi[0] if a else lo[0])
            P = P + f[None, ..., None] * coords[:, v, None, None, None, :]
        return P
    lo, hi = 0.5 * (1.0 - z), 0.5 * (1.0 + z)
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    P = 0.0
    for v, (a, b) in enumerate(corners):
        f = np.einsum("j,i->ji", hi if b else lo, hi if a else lo)
        P = P + f[None, ..., None] * coords[:, v, None, None, :]
    return P


def dirichlet_lift(coords, vertex_ids, nm, fn, nq=None):
    """Global vector of Dirichlet values for the boundary modes of a straight-sided hex
    (coords (E, 8, 3)) or quad (coords (E, 4, 2)) mesh, zero elsewhere: fn(points (n, dim))
    -> values, projected on every boundary face (edge) as the reference projects Dirichlet
    conditions (DisContField.cpp:3630-3663: FwdTrans_BndConstrained of the boundary
    expansion), on the element's own GLL points (nq = nm + 1 by default, the reference's
    TYPE="MODIFIED").  Host side (numpy); pass the result as `lift` to helmsolve."""
    if isinstance(nm, tuple):
        return _dirichlet_lift_aniso(coords, vertex_ids, tuple(int(m) for m in nm), fn, nq)
    if np.ndim(nm) != 0:
        raise MFError("dirichlet_lift: one order for every element (variable order is not "
                      "supported here)")
    coords = np.asarray(coords, dtype=np.float64)
    nq = nm + 1 if nq is None else nq
    shape = "hex" if coords.shape[1] == 8 else "quad"
    z, w, B = _tables(nm, nq)
    l2g, sign, bnd, ng = assembly_map(vertex_ids, nm, shape)
    g = np.zeros(ng)
    P = _points(coords, z)
    E = coords.shape[0]
    dim = 3 if shape == "hex" else 2
    M = nm ** dim
    loc = np.arange(M).reshape((nm,) * dim)  # [r][q][p] or [q][p]
    for e in range(E):
        U = fn(P[e].reshape(-1, coords.shape[2])).reshape((nq,) * dim)
        for d in range(dim):              # normal direction of the face (edge)
            for side in range(2):
                sl = [slice(None)] * dim
                sl[dim - 1 - d] = 0 if side == 0 else nq - 1
                ml = [slice(None)] * dim
                ml[dim - 1 - d] = side
                li = e * M + loc[tuple(ml)].ravel()
                if not np.all(bnd[li]):
                    continue
                F = U[tuple(sl)]
                c = _constrained_2d(F, B, w) if dim == 3 else _constrained_1d(F, B, w)
                g[l2g[li]] = sign[li] * c.ravel()
    return g


# Local faces of a hexahedron in the reference's order (SpatialDomains/HexGeom.cpp): the
# normal direction d (0: xi0, 1: xi1, 2: xi2) and the side (0: xi_d = -1, 1: xi_d = +1).
HEX_FACES = [(2, 0), (1, 0), (0, 1), (1, 1), (0, 0), (2, 1)]


def _hex_map_derivs(coords_e, z):
    """dx/dxi_a (a = 0, 1, 2) of the trilinear map at the GLL grid: (3, nq, nq, nq, 3)."""
    lo, hi = 0.5 * (1.0 - z), 0.5 * (1.0 + z)
    half = np.full_like(z, 0.5)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
               (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    out = np.zeros((3, len(z), len(z), len(z), 3))
    for v, c in enumerate(corners):
        for a in range(3):
            f = [hi if c[d] else lo for d in range(3)]
            f[a] = half if c[a] else -half
            out[a] += np.einsum("k,j,i->kji", f[2], f[1], f[0])[..., None] * coords_e[v]
    return out


def hex_face_quadrature(coords_e, nm, face, nq=None):
    """The quadrature of one face of a straight-sided hexahedron (vertices coords_e (8, 3)
    in local order) on the element's own GLL points, as the reference integrates over the
    face's trace expansion (its points are the element's points in the two tangential
    directions): returns (local mode indices of the face's nm*nm modes, points (nq*nq, 3),
    weights times the surface Jacobian (nq*nq), basis values (nq*nq, nm*nm)).  The other
    element modes vanish on the face."""
    nq = nm + 1 if nq is None else nq
    z, w, B = _tables(nm, nq)
    d, side = HEX_FACES[face]
    t0, t1 = [a for a in range(3) if a != d]
    P = _points(np.asarray(coords_e, dtype=np.float64)[None], z)[0]     # [k][j][i]
    dX = _hex_map_derivs(np.asarray(coords_e, dtype=np.float64), z)
    sl = [slice(None)] * 3
    sl[2 - d] = 0 if side == 0 else nq - 1
    sl = tuple(sl)
    # face arrays indexed [j1][j0] with j0 along t0, j1 along t1 (t0 < t1: t0 is the
    # faster array axis of the [k][j][i] grid)
    pts = P[sl].reshape(-1, 3)
    J = np.linalg.norm(np.cross(dX[t0][sl], dX[t1][sl]), axis=-1).reshape(-1)
    W = np.outer(w, w).reshape(-1) * J
    phi = np.einsum("jb,ia->jiba", B, B).reshape(nq * nq, nm * nm)  # [(j1, j0), (b, a)]
    loc = np.arange(nm ** 3).reshape(nm, nm, nm)                      # [r][q][p]
    ml = [slice(None)] * 3
    ml[2 - d] = side
    modes = loc[tuple(ml)].reshape(-1)                                  # [b][a]
    return modes, pts, W, phi


def boundary_terms(coords, vertex_ids, nm, neumann=(), robin=(), nq=None):
    """The boundary integrals of a straight-sided hex mesh's Neumann and Robin conditions
    (DisContField::v_... / ContField HelmSolve: the boundary expansions' IProductWRTBase of
    the condition's values, and for Robin du/dn + alpha u = g the trace mass matrix times
    alpha that GlobalLinSysIterativeFull::v_DoMatrixMultiply adds through
    AddRobinTraceContribution, GlobalLinSysIterativeFull.cpp:232-252).  neumann: list of
    (faces, g) with faces [(element, local face), ...] and g(points (n, 3)) -> du/dn;
    robin: list of (faces, alpha, g).  Returns (rhs (nglobal,), rows, cols, vals): the
    right-hand-side contribution and the Robin matrix in global coordinates (COO, with the
    signs of the map)."""
    coords = np.asarray(coords, dtype=np.float64)
    l2g, sign, _, ng = assembly_map(vertex_ids, nm, "hex")
    M = nm ** 3
    rhs = np.zeros(ng)
    rows, cols, vals = [], [], []
    for faces, alpha, g in [(f, None, g) for f, g in neumann] + list(robin):
        for e, lf in faces:
            modes, pts, W, phi = hex_face_quadrature(coords[e], nm, lf, nq)
            li = e * M + modes
            gi, s = l2g[li], sign[li].astype(np.float64)
            np.add.at(rhs, gi, s * (phi.T @ (W * g(pts))))
            if alpha is not None:
                Mf = alpha * (phi.T @ (W[:, None] * phi)) * np.outer(s, s)
                rows.append(np.repeat(gi, len(gi)))
                cols.append(np.tile(gi, len(gi)))
                vals.append(Mf.ravel())
    if rows:
        return rhs, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    return rhs, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)


def dirichlet_lift_faces(coords, vertex_ids, nm, fn, faces, nq=None):
    """dirichlet_lift restricted to the listed faces [(element, local face), ...] of a
    straight-sided hex mesh, for conditions that differ by boundary region: returns
    (g, mask), g the global vector of the projected Dirichlet values (zero elsewhere) and
    mask the global modes on those faces' closure (vertices and edges included, which the
    reference also fixes by the Dirichlet condition)."""
    coords = np.asarray(coords, dtype=np.float64)
    nq = nm + 1 if nq is None else nq
    z, w, B = _tables(nm, nq)
    l2g, sign, _, ng = assembly_map(vertex_ids, nm, "hex")
    g = np.zeros(ng)
    mask = np.zeros(ng, dtype=bool)
    P = _points(coords, z)
    M = nm ** 3
    loc = np.arange(M).reshape(nm, nm, nm)
    for e, lf in faces:
        d, side = HEX_FACES[lf]
        U = fn(P[e].reshape(-1, 3)).reshape(nq, nq, nq)
        sl = [slice(None)] * 3
        sl[2 - d] = 0 if side == 0 else nq - 1
        ml = [slice(None)] * 3
        ml[2 - d] = side
        li = e * M + loc[tuple(ml)].ravel()
        c = _constrained_2d(U[tuple(sl)], B, w)
        g[l2g[li]] = sign[li] * c.ravel()
        mask[l2g[li]] = True
    return g, mask


def _dirichlet_lift_aniso(coords, vertex_ids, nm3, fn, nq3=None):
    """dirichlet_lift for hexahedra with nm3[d] modes on nq3[d] GLL points (nm3[d] + 1 by
    default) along direction d, the same for every element: each boundary face is
    projected with the tables of its two in-plane directions."""
    coords = np.asarray(coords, dtype=np.float64)
    nq3 = tuple(m + 1 for m in nm3) if nq3 is None else tuple(nq3)
    tabs = [_tables(nm3[d], nq3[d]) for d in range(3)]
    l2g, sign, bnd, ng = assembly_map(vertex_ids, nm3, "hex")
    g = np.zeros(ng)
    P = _points(coords, [t[0] for t in tabs])
    E = coords.shape[0]
    M = nm3[0] * nm3[1] * nm3[2]
    loc = np.arange(M).reshape(nm3[2], nm3[1], nm3[0])  # [r][q][p]
    for e in range(E):
        U = fn(P[e].reshape(-1, 3)).reshape(nq3[2], nq3[1], nq3[0])
        for d in range(3):                # normal direction of the face
            fast, slow = [x for x in range(3) if x != d]
            for side in range(2):
                sl = [slice(None)] * 3
                sl[2 - d] = 0 if side == 0 else nq3[d] - 1
                ml = [slice(None)] * 3
                ml[2 - d] = side
                li = e * M + loc[tuple(ml)].ravel()
                if not np.all(bnd[li]):
                    continue
                c = _constrained_2d(U[tuple(sl)], tabs[fast][2], tabs[fast][1],
                                    tabs[slow][2], tabs[slow][1])
                g[l2g[li]] = sign[li] * c.ravel()
    return g


class Assembly:
    """Device assembly map of a hex mesh (vertex_ids (E, 8)) or a quad mesh (shape="quad",
    vertex_ids (E, 4)); dirichlet=True fixes every mode on the boundary faces (edges) to
    zero.  nm: the order of every element, per-element orders (variable order; local
    arrays then hold element e's nm[e]^dim coefficients after those of elements < e), or
    for hexahedra orders per direction: a tuple (nm0, nm1, nm2) for every element or an
    (nElmt, 3) array (element e then holds nm0 nm1 nm2 coefficients)."""

    def __init__(self, vertex_ids, nm, dirichlet=True, shape="hex"):
        ids = np.ascontiguousarray(vertex_ids, dtype=np.int64)
        if ids.ndim != 2 or ids.shape[1] != _nverts(shape):
            raise MFError(f"assembly: vertex ids must be (nElmt, {_nverts(shape)}) for {shape}")
        self.shape, self.nElmt, self.nm = shape, ids.shape[0], nm
        dim = 3 if shape == "hex" else 2
        nm3 = _orders3(nm, self.nElmt) if shape == "hex" else None
        nmv = _orders(nm, self.nElmt) if nm3 is None else None
        h = ctypes.c_void_p()
        if nm3 is not None:
            _check(_lib().imf_assembly_create_aniso(
                self.nElmt, nm3.ctypes.data_as(ctypes.c_void_p),
                ids.ctypes.data_as(ctypes.c_void_p), int(dirichlet), ctypes.byref(h)))
            self.nlocal = int(np.sum(np.prod(nm3.astype(np.int64), axis=1)))
        elif nmv is None:
            _check(getattr(_lib(), f"imf_assembly_create_{shape}")(
                self.nElmt, nm, ids.ctypes.data_as(ctypes.c_void_p), int(dirichlet),
                ctypes.byref(h)))
            self.nlocal = self.nElmt * nm ** dim
        else:
            _check(_lib().imf_assembly_create_varp(
                dim, self.nElmt, nmv.ctypes.data_as(ctypes.c_void_p),
                ids.ctypes.data_as(ctypes.c_void_p), int(dirichlet), ctypes.byref(h)))
            self.nlocal = int(np.sum(nmv.astype(np.int64) ** dim))
        self._h = h
        self.nglobal = _lib().imf_assembly_nglobal(h)
        self._robin = None

    def set_dirichlet_mask(self, mask):
        """Replace the Dirichlet modes by mask (nglobal booleans; None: no Dirichlet modes),
        for conditions that differ by boundary region (see dirichlet_lift_faces)."""
        if mask is None:
            _check(_lib().imf_assembly_set_dirichlet_mask(self._h, None))
            return
        m = np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
        if m.shape != (self.nglobal,):
            raise MFError(f"set_dirichlet_mask: expected {self.nglobal} entries, got {m.shape}")
        _check(_lib().imf_assembly_set_dirichlet_mask(self._h, m.ctypes.data_as(ctypes.c_void_p)))

    def set_robin(self, rows, cols, vals):
        """The Robin trace matrix over the global modes (COO, e.g. from boundary_terms),
        added to the solve's operator."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        if not (r.shape == c.shape == v.shape):
            raise MFError("set_robin: rows, cols and vals differ in length")
        _check(_lib().imf_assembly_set_robin(self._h, len(v), r.ctypes.data_as(ctypes.c_void_p),
                                             c.ctypes.data_as(ctypes.c_void_p),
                                             v.ctypes.data_as(ctypes.c_void_p)))
        self._robin = (r.astype(np.int64), c.astype(np.int64), v) if len(v) else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib().imf_assembly_destroy(h)
            self._h = None

    def global_to_local(self, g, out=None, stream=None):
        out = torch.empty(self.nlocal, dtype=torch.float64, device=g.device) if out is None else out
        _check(_lib().imf_global_to_local_device(self._h, ctypes.c_void_p(g.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr()),
                                                 _stream_ptr(stream)))
        return out

    def assemble(self, loc, out=None, stream=None):
        out = torch.empty(self.nglobal, dtype=torch.float64, device=loc.device) if out is None else out
        _check(_lib().imf_assemble_device(self._h, ctypes.c_void_p(loc.data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out

    def apply_local(self, ops, loc):
        """H applied to a local vector by one operator per run of equal-order elements."""
        if not isinstance(ops, (list, tuple)):
            return ops.apply(loc)
        out, o = torch.empty_like(loc), 0
        for op in ops:
            n = int(op.in_size)
            out[o:o + n] = op.apply(loc[o:o + n])
            o += n
        return out

    def helmsolve(self, op, rhs, x=None, tol=1e-9, max_iter=10000, stream=None, lift=None):
        """x with Assemble(H(GlobalToLocal(x))) = rhs on the free modes; returns
        (x, iterations, relative residual).  op: a Helmholtz operator on all elements, or a
        list of operators, one per run of equal-order elements in element order (variable
        order).  Dirichlet modes are zero, or, with `lift` (a global vector holding the
        Dirichlet values, e.g. from dirichlet_lift), equal to lift: the solve is then for
        x - lift with rhs - A lift, as the reference's ContField::HelmSolve treats
        inhomogeneous Dirichlet conditions (A includes the Robin matrix of set_robin)."""
        if lift is not None:
            lift = torch.as_tensor(lift, dtype=torch.float64, device=rhs.device)
            rhs = rhs - self.assemble(self.apply_local(op, self.global_to_local(lift)))
            if self._robin is not None:  # the operator's Robin part acting on the lift
                r, c, v = self._robin
                lh = lift.cpu().numpy()
                rl = np.bincount(r, weights=v * lh[c], minlength=self.nglobal)
                rhs = rhs - torch.from_numpy(rl).to(rhs.device)
        x = torch.empty_like(rhs) if x is None else x
        it, res = ctypes.c_int(), ctypes.c_double()
        ops = list(op) if isinstance(op, (list, tuple)) else [op]
        hs = (ctypes.c_void_p * len(ops))(*[o._h.value for o in ops])
        _check(_lib().imf_helmsolve_cg_groups_device(
            self._h, len(ops), ctypes.cast(hs, ctypes.c_void_p), ctypes.c_void_p(rhs.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), float(tol), int(max_iter), ctypes.byref(it),
            ctypes.byref(res), _stream_ptr(stream)))
        if lift is not None:
            x += lift
        return x, it.value, res.value
