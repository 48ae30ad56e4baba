"""Variable-order C0 numbering (host part of the library) checked on the CPU.

The reference's rule for modal expansions of different order (AssemblyMapCG.cpp:1380-1440,
1771-1791): a shared edge carries the fewest interior modes of the elements around it, a
shared face the fewest of its two elements, and an element's modes beyond those get
global id 0 and sign 0.  What pins it here:

* with one order everywhere the variable-order map is the uniform map, bit for bit;
* a known answer counted by hand: two hexes of 4 and 6 modes sharing a face have
  12 vertex + 56 edge + 104 face + 72 interior = 244 global modes (a max rule would give
  a different count);
* continuity: a random global vector, scattered with the map and evaluated with each
  element's own basis on a common set of points, takes one value at every physical point
  shared by several elements, with random per-element orders and every element in a random
  rotation of its reference frame (a surplus mode left in place breaks this by O(1));
* every global mode is reached by some local mode with a non-zero sign;
* the assembled Helmholtz solution on a mesh with mixed orders lies between the uniform
  lowest- and highest-order solutions in accuracy (the discrete spaces are nested).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import dense_ref as DR
import oracle as O
from ithaca_hip import assembly, mesh
from test_assembly_cpu import _mesh, _qmesh, manufactured


def _offsets(nmv, dim):
    return np.concatenate([[0], np.cumsum(np.asarray(nmv, dtype=np.int64) ** dim)])


@pytest.mark.parametrize("shape,nm", [("hex", 4), ("hex", 6), ("quad", 5)])
def test_uniform_orders_give_the_uniform_map(shape, nm):
    verts, ids = (_mesh if shape == "hex" else _qmesh)(3, True, seed=4)
    a = assembly.assembly_map(ids, nm, shape)
    b = assembly.assembly_map(ids, np.full(len(ids), nm), shape)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3] == b[3]


def test_two_hexes_known_count():
    # unit cubes [0,1]^3 and [1,2]x[0,1]^2 sharing the face x = 1; vertex id = grid index
    def vid(ix, iy, iz):
        return (iz * 2 + iy) * 3 + ix
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
               (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    ids = np.array([[vid(x0 + cx, cy, cz) for cx, cy, cz in corners] for x0 in (0, 1)])
    l2g, sign, bnd, ng = assembly.assembly_map(ids, [4, 6], "hex")
    assert ng == 12 + (4 * 2 + 8 * 2 + 8 * 4) + (4 + 5 * 4 + 5 * 16) + (8 + 64)
    assert set(np.unique(l2g[sign != 0])) == set(range(ng))
    # the 6-mode element drops, on the shared face x = 1 (its p = 0 side), every face mode
    # with an index >= 4 and every edge mode >= 4 on the face's 4 edges
    e1 = sign[64:].reshape(6, 6, 6)  # [r][q][p]
    face = e1[:, :, 0]
    zero = np.zeros((6, 6), dtype=bool)
    zero[4:, :] = zero[:, 4:] = True
    assert np.array_equal(face == 0, zero)
    assert np.all(sign[:64] != 0) and np.all(e1[:, :, 1:] != 0)
    # dropped modes are not boundary modes; the shared face's interior is not on the boundary
    assert np.all(bnd[sign == 0] == 0)
    assert np.all(bnd[64:].reshape(6, 6, 6)[2:4, 2:4, 0] == 0)
    assert bnd[sign != 0].sum() > 0


def _continuity(shape, verts, ids, nmv, nq):
    dim = 3 if shape == "hex" else 2
    l2g, sign, bnd, ng = assembly.assembly_map(ids, nmv, shape)
    assert set(np.unique(l2g[sign != 0])) == set(range(ng))
    assert (sign == 0).any() and (sign < 0).any()
    off = _offsets(nmv, dim)
    x = np.random.default_rng(1).uniform(-1, 1, ng)
    pts, vals = [], []
    for e in range(len(ids)):
        sl = slice(off[e], off[e + 1])
        m = DR.element_mats(shape, int(nmv[e]), nq, verts[e])
        pts.append(m["X"])
        vals.append(m["B"] @ (sign[sl] * x[l2g[sl]]))
    P, V = np.concatenate(pts), np.concatenate(vals)
    _, inv = np.unique(np.round(P * 1e9).astype(np.int64), axis=0, return_inverse=True)
    inv = inv.ravel()
    lo = np.full(inv.max() + 1, np.inf)
    hi = np.full(inv.max() + 1, -np.inf)
    np.minimum.at(lo, inv, V)
    np.maximum.at(hi, inv, V)
    shared = np.bincount(inv) > 1
    assert shared.sum() > 0
    return np.max(hi[shared] - lo[shared])


def test_variable_order_continuity_hex():
    verts, ids = _mesh(3, True, seed=11, jitter=0.1)
    nmv = np.random.default_rng(5).integers(3, 7, len(ids))
    assert _continuity("hex", verts, ids, nmv, 8) < 1e-12


def test_variable_order_continuity_quad():
    verts, ids = _qmesh(4, True, seed=12, jitter=0.1)
    nmv = np.random.default_rng(6).integers(3, 9, len(ids))
    assert _continuity("quad", verts, ids, nmv, 10) < 1e-12


def varp_system(verts, ids, nmv, lam, f_exact, shape="hex"):
    """Global CSR matrix and rhs (Dirichlet modes masked) from dense element matrices with
    each element's own order and nq = nm + 1 points."""
    dim = 3 if shape == "hex" else 2
    l2g, sign, bnd, ng = assembly.assembly_map(ids, nmv, shape)
    off = _offsets(nmv, dim)
    rows, cols, vals = [], [], []
    rhs = np.zeros(ng)
    for e in range(len(ids)):
        nm = int(nmv[e])
        m = DR.element_mats(shape, nm, nm + 1, verts[e])
        He = DR.helmholtz_matrix(m, lam)
        sl = slice(off[e], off[e + 1])
        g, s = l2g[sl], sign[sl].astype(float)
        Mloc = len(g)
        rows.append(np.repeat(g, Mloc))
        cols.append(np.tile(g, Mloc))
        vals.append((He * np.outer(s, s)).ravel())
        np.add.at(rhs, g, s * (m["B"].T @ (m["W"] * m["J"] * f_exact(m["X"]))))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ng, ng))
    dm = np.zeros(ng, dtype=bool)
    dm[np.unique(l2g[bnd == 1])] = True
    return A, rhs, dm, l2g, sign


def varp_solve(verts, ids, nmv, lam, shape="hex"):
    u, f = manufactured(lam, 3 if shape == "hex" else 2)
    A, rhs, dm, l2g, sign = varp_system(verts, ids, nmv, lam, f, shape)
    free = ~dm
    x = np.zeros(len(rhs))
    x[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free])
    return x, l2g, sign, u


def varp_error(verts, ids, nmv, x, l2g, sign, u, shape="hex"):
    dim = 3 if shape == "hex" else 2
    off = _offsets(nmv, dim)
    num = den = 0.0
    for e in range(len(ids)):
        nm = int(nmv[e])
        m = DR.element_mats(shape, nm, nm + 1, verts[e])
        sl = slice(off[e], off[e + 1])
        ue = m["B"] @ (sign[sl] * x[l2g[sl]])
        ex = u(m["X"])
        num += np.sum(m["W"] * m["J"] * (ue - ex) ** 2)
        den += np.sum(m["W"] * m["J"] * ex ** 2)
    return np.sqrt(num / den)


@pytest.mark.parametrize("shape", ["hex", "quad"])
def test_mixed_orders_between_uniform_orders(shape):
    if shape == "hex":
        verts, ids = _mesh(2, True, seed=3, jitter=0.05)
        lo, hi = 4, 6
    else:
        verts, ids = _qmesh(4, True, seed=3, jitter=0.05)
        lo, hi = 3, 6
    E = len(ids)
    nmv = np.where(np.arange(E) % 2 == 0, lo, hi)
    errs = {}
    for name, orders in (("low", np.full(E, lo)), ("mixed", nmv), ("high", np.full(E, hi))):
        x, l2g, sign, u = varp_solve(verts, ids, orders, 10.0, shape)
        errs[name] = varp_error(verts, ids, orders, x, l2g, sign, u, shape)
    print(shape, errs)
    assert errs["high"] < errs["mixed"] < errs["low"]
