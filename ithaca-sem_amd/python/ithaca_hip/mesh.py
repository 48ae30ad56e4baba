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
