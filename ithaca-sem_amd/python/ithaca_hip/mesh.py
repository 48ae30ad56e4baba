"""Synthetic meshes and inputs of the measured workloads.

Elements are ordered lexicographically (x fastest); each element lists its vertices in
the geometric order of the reference hex / quad (StdHexExp.cpp:1042-1082): vertex v has
xi0 = +1 for v%4 in {1,2}, xi1 = +1 for v%4 in {2,3}, xi2 = +1 for v > 3.
Random numbers come from ``numpy.random.default_rng(seed)`` (PCG64).
"""
import numpy as np

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
QUAD_CORNERS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])


def hex_box(n, jitter=0.0, rng=None):
    """n^3 hexes on [0,1]^3; interior vertices moved by U[-jitter, jitter]*h per coordinate.

    Returns verts (n^3, 8, 3)."""
    h = 1.0 / n
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1)  # [ix, iy, iz, 3]
    if jitter > 0.0:
        d = rng.uniform(-jitter, jitter, size=(n - 1, n - 1, n - 1, 3)) * h
        X[1:-1, 1:-1, 1:-1, :] += d
    ez, ey, ex = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    verts = np.empty((n ** 3, 8, 3))
    for v, (cx, cy, cz) in enumerate(HEX_CORNERS):
        verts[:, v, :] = X[ex + cx, ey + cy, ez + cz]
    return verts


def quad_box(n, jitter=0.0, rng=None):
    """n^2 quads on [0,1]^2 (optionally jittered interior vertices). verts (n^2, 4, 2)."""
    h = 1.0 / n
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1)
    if jitter > 0.0:
        X[1:-1, 1:-1, :] += rng.uniform(-jitter, jitter, size=(n - 1, n - 1, 2)) * h
    ey, ex = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ex, ey = ex.ravel(), ey.ravel()
    verts = np.empty((n ** 2, 4, 2))
    for v, (cx, cy) in enumerate(QUAD_CORNERS):
        verts[:, v, :] = X[ex + cx, ey + cy]
    return verts


def workload_quad_helmholtz():
    """32x32 regular quads on [0,1]^2, P=4 (5 modes), 6 GLL points, lambda = 1,
    coefficients U[-1,1] from seed 0."""
    nm, nq, n = 5, 6, 32
    rng = np.random.default_rng(0)
    verts = quad_box(n)
    u = rng.uniform(-1.0, 1.0, size=n * n * nm * nm)
    return dict(shape="quad", nm=nm, nq=nq, nElmt=n * n, verts=verts, u=u, lam=1.0,
                deformed=False)


def workload_hex_helmholtz(n=48):
    """n^3 hexes on [0,1]^3 with interior vertices jittered by U[-0.05,0.05]*h (seed 1),
    P=5 (6 modes), 7 GLL points, lambda = 10, coefficients U[-1,1] (same stream)."""
    nm, nq = 6, 7
    rng = np.random.default_rng(1)
    verts = hex_box(n, 0.05, rng)
    u = rng.uniform(-1.0, 1.0, size=n ** 3 * nm ** 3)
    return dict(shape="hex", nm=nm, nq=nq, nElmt=n ** 3, verts=verts, u=u, lam=10.0,
                deformed=True)


def hex_quad_points(verts, z):
    """Physical coordinates (E, Q, 3) of the nq^3 GLL points of straight-sided hexes under
    the trilinear map of their vertices (torch, on verts' device); point index
    (k*nq + j)*nq + i with xi = (z_i, z_j, z_k)."""
    import torch
    dev = verts.device
    zt = torch.as_tensor(np.asarray(z), dtype=torch.float64, device=dev)
    k, j, i = torch.meshgrid(zt, zt, zt, indexing="ij")
    xi = torch.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], dim=1)  # (Q, 3)
    c = torch.as_tensor(HEX_CORNERS, dtype=torch.float64, device=dev)      # (8, 3)
    shp = torch.where(c[None] > 0, (1.0 + xi[:, None]) / 2, (1.0 - xi[:, None]) / 2)
    N = torch.prod(shp, dim=2)                                               # (Q, 8)
    E = verts.shape[0]
    X = N @ verts.permute(1, 0, 2).reshape(8, E * 3)                         # (Q, E*3)
    return X.reshape(-1, E, 3).permute(1, 0, 2).contiguous()


def quad_points(verts, z):
    """Physical coordinates (E, Q, 2) of the nq^2 GLL points of straight-sided quads under
    the bilinear map of their vertices (torch, on verts' device); point index j*nq + i."""
    import torch
    dev = verts.device
    zt = torch.as_tensor(np.asarray(z), dtype=torch.float64, device=dev)
    j, i = torch.meshgrid(zt, zt, indexing="ij")
    xi = torch.stack([i.reshape(-1), j.reshape(-1)], dim=1)                  # (Q, 2)
    c = torch.as_tensor(QUAD_CORNERS, dtype=torch.float64, device=dev)      # (4, 2)
    shp = torch.where(c[None] > 0, (1.0 + xi[:, None]) / 2, (1.0 - xi[:, None]) / 2)
    N = torch.prod(shp, dim=2)                                               # (Q, 4)
    E = verts.shape[0]
    X = N @ verts.permute(1, 0, 2).reshape(4, E * 2)                         # (Q, E*2)
    return X.reshape(-1, E, 2).permute(1, 0, 2).contiguous()


def pod_fields(n_fields=20, n_snap=512, seed=2):
    """Snapshot recipe of the POD workload: field k is prod_i sin(m_{k,i} pi x_i) with
    integer wavenumbers m ~ U{1..6}; snapshot j is sum_k a_{jk} field_k with
    a_{jk} ~ N(0,1) 2^{-(k+1)} (k = 0..n_fields-1).  Returns (m (n_fields, 3),
    a (n_snap, n_fields)) from numpy PCG64(seed)."""
    rng = np.random.default_rng(seed)
    m = rng.integers(1, 7, size=(n_fields, 3))
    a = rng.standard_normal((n_snap, n_fields)) * 2.0 ** -np.arange(1, n_fields + 1)
    return m, a


def hex_box_ids(n):
    """Global vertex ids (n^3, 8) of hex_box(n): vertex (ix, iy, iz) of the (n+1)^3 grid has
    id (iz*(n+1) + iy)*(n+1) + ix; same element and vertex order as hex_box."""
    ez, ey, ex = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    ids = np.empty((n ** 3, 8), dtype=np.int64)
    for v, (cx, cy, cz) in enumerate(HEX_CORNERS):
        ids[:, v] = ((ez + cz) * (n + 1) + (ey + cy)) * (n + 1) + (ex + cx)
    return ids


def quad_box_ids(n):
    """Global vertex ids (n^2, 4) of quad_box(n): vertex (ix, iy) of the (n+1)^2 grid has id
    iy*(n+1) + ix; same element and vertex order as quad_box."""
    ey, ex = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ex, ey = ex.ravel(), ey.ravel()
    ids = np.empty((n ** 2, 4), dtype=np.int64)
    for v, (cx, cy) in enumerate(QUAD_CORNERS):
        ids[:, v] = (ey + cy) * (n + 1) + (ex + cx)
    return ids


def rotate_quads(verts, ids, rng):
    """Each quad re-listed from a random starting vertex (the 4 proper rotations of the
    reference square): same physical element, different local orientation."""
    shift = rng.integers(0, 4, size=len(verts))
    p = (np.arange(4)[None, :] + shift[:, None]) % 4
    return (np.take_along_axis(verts, p[:, :, None], axis=1),
            np.take_along_axis(ids, p, axis=1))


def cube_rotations():
    """The 24 proper rotations of the reference cube as vertex permutations: element vertex
    k of the rotated frame is vertex perm[k] of the original frame."""
    import itertools
    corners = HEX_CORNERS * 2 - 1
    index = {tuple(c): v for v, c in enumerate(corners)}
    perms = []
    for axes in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=int)
            for i, (a, s) in enumerate(zip(axes, signs)):
                R[i, a] = s
            if round(np.linalg.det(R)) != 1:
                continue
            perms.append([index[tuple(R.T @ c)] for c in corners])
    return np.array(perms)


def rotate_elements(verts, ids, rng):
    """Each element re-listed in a random proper rotation of its reference frame (same
    physical element, different local orientation)."""
    rots = cube_rotations()
    pick = rng.integers(0, len(rots), size=len(verts))
    p = rots[pick]
    return (np.take_along_axis(verts, p[:, :, None], axis=1),
            np.take_along_axis(ids, p, axis=1))
