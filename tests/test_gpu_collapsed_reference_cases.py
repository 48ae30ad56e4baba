"""The reference's MatrixFree IProductWRTDerivBase unit tests for the collapsed shapes, on
the GPU: the elements, orders, point sets and fields of
UnitTests/Collections/TestTriCollection.cpp:1956-, :2133- (TestTriIProductWRTDerivBase_
MatrixFree_UniformP_{Undeformed,Deformed}), TestTetCollection.cpp:2537-,
TestPrismCollection.cpp (TestPrismIProductWRTDerivBase_MatriFree_UniformP_{Undeformed,
Deformed,Deformed_OverInt}_MultiElmt) and TestPyrCollection.cpp:2084-
(TestPyrIProductWRTDerivBase_MatrixFree_Deformed_MultiElmt).  Those tests compare the
MatrixFree operator with the element's own IProductWRTDerivBase summed over directions to
BOOST_CHECK_CLOSE 1e-8 (percent); here the GPU operator is compared with the same sum
formed from independent dense gradient matrices (tests/dense_ref.py), to 1e-10
relative, the element repeated over several elements as in the MultiElmt cases."""
import numpy as np
import pytest
import torch

import dense_ref as R
import ithaca_hip as H

pytestmark = pytest.mark.gpu

E = 10

CASES = [
    # name, shape, vertices, nm, (nq0, nq1[, nq2]), affine
    ("tri_undeformed", "tri", [(-1, -1), (1, -1), (-1, 1)], 4, (5, 4), True),
    ("tri_deformed", "tri", [(-1, -1), (1, -1), (-1, 2)], 4, (5, 4), True),
    ("tet_undeformed", "tet", [(-2, -3, -4), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)], 4,
     (5, 4, 4), True),
    ("prism_undeformed", "prism", [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                   (-1, -1, 1), (-1, 1, 1)], 6, (7, 7, 6), True),
    ("prism_deformed", "prism", [(-2, -3, -4), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                 (-1, -1, 1), (-1, 1, 1)], 6, (7, 7, 6), False),
    ("prism_deformed_overint", "prism", [(-2, -3, -4), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                                         (-1, -1, 1), (-1, 1, 1)], 6, (12, 12, 11), False),
    ("pyr_deformed", "pyr", [(-1.5, -1.5, -1.5), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                             (-1, -1, 1)], 4, (5, 5, 4), False),
]


def _mats(shape, nm, nq, verts):
    if shape == "tri":
        return R.tri_element_mats(nm, nq[0], nq[1], verts)
    if shape == "tet":
        return R.tet_element_mats(nm, nq[0], nq[1], nq[2], verts)
    if shape == "prism":
        return R.prism_element_mats(nm, nq[0], nq[2], verts)
    return R.pyr_element_mats(nm, nq[0], nq[1], nq[2], verts)


def _fields(X):
    """The unit tests' fields at the element's points (X: dim x Q coordinates)."""
    if X.shape[0] == 2:
        x, y = X
        return [np.sin(x) * np.cos(y), np.cos(x) * np.sin(y)]
    x, y, z = X
    return [np.sin(x) * np.cos(y) * np.sin(z), np.cos(x) * np.sin(y) * np.cos(z),
            np.cos(x) * np.sin(y) * np.sin(z)]


@pytest.mark.parametrize("name,shape,verts,nm,nq,affine", CASES, ids=[c[0] for c in CASES])
def test_reference_iproduct_wrt_deriv_base_cases(name, shape, verts, nm, nq, affine):
    print(f"[case] {name}: nm={nm} nq={nq}", flush=True)
    v = np.asarray(verts, dtype=np.float64)
    m = _mats(shape, nm, nq, v)
    dim = v.shape[1]
    f = _fields(m["X"])
    WJ = m["W"] * m["J"]
    ref = sum(m["Gphys"][c].T @ (WJ * f[c]) for c in range(dim))
    key = f"IProductWRTDerivBase_{shape.capitalize()}_{'Regular' if affine else 'Deformed'}"
    op = H.Operator(key, nm, nq, E)
    vd = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(v, (E,) + v.shape))).cuda()
    jac, df = H.geom_device(shape, vd, nq, None, not affine)
    op.set_jac(jac)
    op.set_df(df)
    inp = torch.from_numpy(np.concatenate([np.tile(fc, E) for fc in f])).cuda()
    got = op.apply(inp).cpu().numpy().reshape(E, -1)
    for e in range(E):
        assert np.linalg.norm(got[e] - ref) < 1e-10 * np.linalg.norm(ref), (name, e)
