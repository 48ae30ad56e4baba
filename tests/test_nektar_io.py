"""Nektar++ XML mesh and .fld readers (host only) on the reference's own FieldConvert
fixtures (tests/golden/nektar) and on small synthetic files.

What pins the readers:
* the local vertex order restated from QuadGeom/HexGeom gives right-handed elements;
* the field coefficients, read in that local frame and evaluated with the dense
  Modified_A basis of tests/dense_ref.py, form a continuous field (a wrong frame or a
  wrong data layout breaks continuity across element boundaries by O(1));
* the hex fixture's fields are the session's initial conditions
  (u = sin 2 pi y + cos 6 pi y, v and w the same in x and z): the field read as u
  correlates with the y function only.
"""
import os

import numpy as np
import pytest

import dense_ref as DR
from ithaca_hip import nektar_io as N

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nektar")


def _fields_at_points(shape, nm, nq, X, coefs):
    pts, vals = [], []
    M = nm ** (3 if shape == "hex" else 2)
    for e in range(len(X)):
        m = DR.element_mats(shape, nm, nq, X[e])
        assert m["J"].min() > 0
        pts.append(m["X"])
        vals.append(m["B"] @ coefs[e * M:(e + 1) * M])
    return np.concatenate(pts), np.concatenate(vals)


def _max_jump(P, V):
    _, inv = np.unique(np.round(P * 1e9).astype(np.int64), axis=0, return_inverse=True)
    inv = inv.ravel()
    lo = np.full(inv.max() + 1, np.inf)
    hi = np.full(inv.max() + 1, -np.inf)
    np.minimum.at(lo, inv, V)
    np.maximum.at(hi, inv, V)
    shared = np.bincount(inv) > 1
    assert shared.sum() > 0
    return np.max(hi[shared] - lo[shared])


def _right_handed(X, dim):
    """Signed Jacobian of the multilinear map at the reference centre, per element."""
    if dim == 3:
        a = (X[:, 1] - X[:, 0] + X[:, 2] - X[:, 3] + X[:, 5] - X[:, 4] + X[:, 6] - X[:, 7])
        b = (X[:, 3] - X[:, 0] + X[:, 2] - X[:, 1] + X[:, 7] - X[:, 4] + X[:, 6] - X[:, 5])
        c = X[:, 4:].sum(1) - X[:, :4].sum(1)
        return np.einsum("ei,ei->e", a, np.cross(b, c))
    a = X[:, 1] - X[:, 0] + X[:, 2] - X[:, 3]
    b = X[:, 3] - X[:, 0] + X[:, 2] - X[:, 1]
    return a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]


def test_hex_mesh_and_fields():
    m = N.read_mesh(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml"))
    assert m.dim == 3
    assert (len(m.coords), len(m.edges), len(m.faces), len(m.elements)) == (27, 54, 36, 8)
    assert m.expansions[0]["nummodes"] == 4 and m.expansions[0]["type"] == "MODIFIED"
    blocks = N.read_fld(os.path.join(HERE, "Hex_channel_C0helmsmoothing.fld"))
    assert len(blocks) == 1
    b = blocks[0]
    assert b.fields == ["u", "v", "w", "p"] and b.nummodes == [4, 4, 4]
    assert list(b.ids) == list(range(8))
    X = m.element_coords(b.ids)
    assert np.all(_right_handed(X, 3) > 0)
    # every vertex of the unit cube's 3^3 grid, each element an axis-aligned 0.5-cube
    assert np.allclose(X.max(1) - X.min(1), 0.5)
    nm, nq = 4, 5
    f1 = lambda s: np.sin(2 * np.pi * s) + np.cos(6 * np.pi * s)
    for name, axis in (("u", 1), ("v", 0), ("w", 2)):
        P, V = _fields_at_points("hex", nm, nq, X, N.field_coefficients(blocks, name, b.ids))
        assert _max_jump(P, V) < 1e-12
        assert np.corrcoef(V, f1(P[:, axis]))[0, 1] > 0.99
        for other in {0, 1, 2} - {axis}:
            assert abs(np.corrcoef(V, f1(P[:, other]))[0, 1]) < 1e-6
    # p = u + v + w (linear in the data)
    cs = [N.field_coefficients(blocks, f, b.ids) for f in "uvwp"]
    assert np.allclose(cs[0] + cs[1] + cs[2], cs[3], atol=1e-10)


def test_quad_mesh_with_offset_ids_and_fields():
    m = N.read_mesh(os.path.join(HERE, "interpfield_4quad.xml"))
    blocks = N.read_fld(os.path.join(HERE, "interpfield_4quad.fld"))
    b = blocks[0]
    assert m.dim == 2 and list(b.ids) == [1655, 1656, 1657, 1658]
    assert b.fields == ["rho"] and b.nummodes == [5, 5]
    X = m.element_coords(b.ids)
    assert np.all(_right_handed(X, 2) > 0)
    ids = m.vertex_ids(b.ids)
    assert len(np.unique(ids)) == 10  # a strip of 4 quads
    P, V = _fields_at_points("quad", 5, 6, X, N.field_coefficients(blocks, "rho", b.ids))
    assert _max_jump(P, V) < 1e-9 * np.abs(V).max()


def test_text_format_mesh_and_field(tmp_path):
    """Uncompressed entities and data: 2 quads, listed with edges in both directions."""
    mesh = """<?xml version="1.0"?>
<NEKTAR><GEOMETRY DIM="2" SPACE="2">
<VERTEX><V ID="0">0 0 0</V><V ID="1">1 0 0</V><V ID="2">2 0 0</V>
<V ID="3">0 1 0</V><V ID="4">1 1 0</V><V ID="5">2 1 0</V></VERTEX>
<EDGE><E ID="0">0 1</E><E ID="1">1 4</E><E ID="2">4 3</E><E ID="3">3 0</E>
<E ID="4">2 1</E><E ID="5">5 2</E><E ID="6">4 5</E></EDGE>
<ELEMENT><Q ID="0">0 1 2 3</Q><Q ID="7">4 5 6 1</Q></ELEMENT>
<COMPOSITE><C ID="0"> Q[0,7] </C></COMPOSITE><DOMAIN> C[0] </DOMAIN>
</GEOMETRY>
<EXPANSIONS><E COMPOSITE="C[0]" NUMMODES="2" FIELDS="u" TYPE="MODIFIED" /></EXPANSIONS>
</NEKTAR>"""
    p = tmp_path / "m.xml"
    p.write_text(mesh)
    m = N.read_mesh(str(p))
    assert list(m.elements) == [0, 7]
    assert m.vertex_ids().tolist() == [[0, 1, 4, 3], [1, 2, 5, 4]]
    assert np.all(_right_handed(m.element_coords(), 2) > 0)
    # u = x + 2y with 2 modes (the vertex modes carry the vertex values)
    X = m.element_coords()
    coef = np.concatenate([X[e, :, 0] + 2 * X[e, :, 1] for e in range(2)])
    # local mode order p + 2q: (0,0) (1,0) (0,1) (1,1) = vertices 0, 1, 3, 2
    coef = coef.reshape(2, 4)[:, [0, 1, 3, 2]].ravel()
    fld = tmp_path / "f.fld"
    fld.write_text('<?xml version="1.0"?><NEKTAR><ELEMENTS FIELDS="u" SHAPE="Quadrilateral" '
                   'BASIS="Modified_A,Modified_A" NUMMODESPERDIR="UNIORDER:2,2" ID="0,7">'
                   + " ".join(repr(float(v)) for v in coef) + "</ELEMENTS></NEKTAR>")
    blocks = N.read_fld(str(fld))
    P, V = _fields_at_points("quad", 2, 3, X, N.field_coefficients(blocks, "u", [0, 7]))
    assert np.allclose(V, P[:, 0] + 2 * P[:, 1], atol=1e-14)


@pytest.mark.parametrize("attrs,msg", [
    ('FIELDS="u" SHAPE="Triangle" BASIS="Modified_A,Modified_B" NUMMODESPERDIR="UNIORDER:2,2" ID="0"', "shape"),
    ('FIELDS="u" SHAPE="Quadrilateral" BASIS="Modified_A,Modified_A" NUMMODESPERDIR="MIXORDER:2,2,3" ID="0"', "MIXORDER"),
    ('FIELDS="u" SHAPE="Quadrilateral" BASIS="Modified_A,Modified_A" NUMMODESPERDIR="MIXORDER:2,2" ID="0-1"', "MIXORDER"),
    ('FIELDS="u" SHAPE="Quadrilateral" BASIS="Modified_A,Modified_A" NUMMODESPERDIR="UNIORDER:2,2" ID=""', "element ids"),
    ('FIELDS="u" SHAPE="Quadrilateral" BASIS="Modified_A,Modified_A" NUMMODESPERDIR="UNIORDER:2,2" ID="0-1"', "length"),
])
def test_fld_rejects_what_it_cannot_read(tmp_path, attrs, msg):
    f = tmp_path / "bad.fld"
    f.write_text(f"<NEKTAR><ELEMENTS {attrs}>1 2 3 4</ELEMENTS></NEKTAR>")
    with pytest.raises(N.NektarIOError, match=msg):
        N.read_fld(str(f))


def test_mesh_rejects_curved_elements_and_unknown_ids(tmp_path):
    base = os.path.join(HERE, "interpfield_4quad.xml")
    text = open(base).read()
    curved = text.replace("</GEOMETRY>", '<CURVED><E ID="0" EDGEID="0" TYPE="PolyEvenlySpaced" '
                                         'NUMPOINTS="3">0 0 0 0.5 0.1 0 1 0 0</E></CURVED></GEOMETRY>')
    p = tmp_path / "c.xml"
    p.write_text(curved)
    with pytest.raises(N.NektarIOError, match="curved"):
        N.read_mesh(str(p))
    with pytest.raises(N.NektarIOError, match="not in the mesh"):
        N.read_mesh(base).element_coords([1])


def test_expansion_arrays():
    m = N.read_mesh(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml"))
    blocks = N.read_fld(os.path.join(HERE, "Hex_channel_C0helmsmoothing.fld"))
    X, ids, c, nm, nq = N.expansion_arrays(m, blocks, "p")
    assert X.shape == (8, 8, 3) and ids.shape == (8, 8) and c.shape == (8 * 64,)
    assert (nm, nq) == (4, 5)
    with pytest.raises(N.NektarIOError, match="no data"):
        N.expansion_arrays(m, blocks, "T")


# ---------------------------------------------------------------- writing

def _elements(path):
    import re
    with open(path) as fh:
        m = re.search(r"<ELEMENTS\b([^>]*)>(.*?)</ELEMENTS>", fh.read(), re.S)
    return m.group(1).strip(), "".join(m.group(2).split())


@pytest.mark.parametrize("name,shape", [("Hex_channel_C0helmsmoothing", "hex"),
                                        ("interpfield_4quad", "quad")])
def test_writer_reproduces_the_reference_files(tmp_path, name, shape):
    """The reference's own .fld files, read and written back: the ELEMENTS attributes and
    the compressed data text are identical to the reference's, byte for byte (zlib at the
    default level, zero padding to a multiple of 3, base64, the id range syntax)."""
    src = os.path.join(HERE, name + ".fld")
    b = N.read_fld(src)[0]
    out = tmp_path / "w.fld"
    N.write_fld(str(out), {f: b.data[f] for f in b.fields}, b.ids, b.nummodes[0], shape,
                metadata={"Time": 0})
    assert _elements(str(out)) == _elements(src)
    c = N.read_fld(str(out))[0]
    assert c.fields == b.fields and list(c.ids) == list(b.ids)
    for f in b.fields:
        assert np.array_equal(c.data[f], b.data[f])


def test_writer_text_round_trip_and_metadata(tmp_path):
    rng = np.random.default_rng(0)
    ids = [3, 4, 5, 9, 2]
    u, v = rng.standard_normal(5 * 27), rng.standard_normal(5 * 27)
    out = tmp_path / "t.fld"
    N.write_fld(str(out), {"u": u, "v": v}, ids, 3, "hex", metadata={"Time": 1.5, "Kinvis": 0.01},
                compressed=False)
    b = N.read_fld(str(out))[0]
    assert list(b.ids) == ids and b.nummodes == [3, 3, 3]
    assert np.array_equal(N.field_coefficients([b], "u", ids), u)
    assert np.array_equal(N.field_coefficients([b], "v", [9]), v[3 * 27:4 * 27])
    md = N.read_fld_metadata(str(out))
    assert md["Time"] == "1.5" and md["Kinvis"] == "0.01" and "Provenance/Timestamp" in md
    ref = N.read_fld_metadata(os.path.join(HERE, "Hex_channel_C0helmsmoothing.fld"))
    assert ref["Provenance/NektarVersion"] == "5.0.0" and ref["Time"] == "0"
    assert ref["SessionName0"] == "Hex_channel_m3.xml"


def test_seq_string_inverts_the_range_syntax():
    # the example of ParseUtils::GenerateSeqVector's documentation, and edge cases
    seq = [1, 2, 3, 4, 6, 7, 8, 5, 2, 3]
    s = N.seq_string(seq)
    assert s == "1-4,6-8,5,2-3"
    assert list(N._id_list(s)) == seq
    assert N.seq_string([7]) == "7" and N.seq_string([1, 1, 2]) == "1-2"
    assert N.seq_string([]) == ""


def test_writer_rejects_bad_input(tmp_path):
    out = str(tmp_path / "x.fld")
    with pytest.raises(N.NektarIOError, match="expected"):
        N.write_fld(out, {"u": np.zeros(10)}, [0], 2, "hex")
    with pytest.raises(N.NektarIOError, match="shape"):
        N.write_fld(out, {"u": np.zeros(8)}, [0], 2, "tet")
    with pytest.raises(N.NektarIOError, match="XML names"):
        N.write_fld(out, {"u v": np.zeros(8)}, [0], 2, "hex")


def test_mixorder_file_of_the_reference(tmp_path):
    """A MIXORDER block of the reference (surface fields of a variable-order run, quads of 2
    and 6 modes): the data split by each element's counts fills the block exactly, and the
    block written back is byte-identical to the reference's (NUMMODESPERDIR list, ids,
    compressed data)."""
    src = os.path.join(HERE, "naca0012_3D_bnd_P0000015.fld")
    b = N.read_fld(src)[0]
    assert b.nummodes is None and b.shape == "Quadrilateral"
    assert set(np.unique(b.modes)) == {2, 6} and np.all(b.modes[:, 0] == b.modes[:, 1])
    assert len(b.fields) == 7
    for f in b.fields:
        assert [len(r) for r in b.data[f]] == list(b.modes.prod(axis=1))
    # Shear_mag is the magnitude of the shear vector at the vertex modes (which carry the
    # vertex values of the modified basis)
    # vertex modes (p, q) in {0, 1}^2 at p + q nm
    vm = [np.array([0, 1, n, n + 1]) for n in b.modes[:, 0]]
    sx, sy, sz, sm = (np.array([r[v] for r, v in zip(b.data[f], vm)]) for f in
                      ("Shear_x", "Shear_y", "Shear_z", "Shear_mag"))
    assert np.allclose(sm, np.sqrt(sx ** 2 + sy ** 2 + sz ** 2), rtol=1e-8, atol=1e-12)
    out = tmp_path / "m.fld"
    N.write_fld(str(out), {f: b.data[f] for f in b.fields}, b.ids, b.modes[:, 0], "quad")
    assert _elements(str(out)) == _elements(src)
    c = N.read_fld(str(out))[0]
    for f in b.fields:
        assert all(np.array_equal(x, y) for x, y in zip(c.data[f], b.data[f]))


def test_mixorder_round_trip_into_variable_order_arrays(tmp_path):
    """A hex field of mixed order written and read back gives per-element orders, which the
    variable-order assembly takes as they are."""
    m = N.read_mesh(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml"))
    nmv = np.array([3, 4, 5, 3, 4, 5, 6, 2])
    rng = np.random.default_rng(2)
    u = [rng.standard_normal(n ** 3) for n in nmv]
    out = tmp_path / "v.fld"
    N.write_fld(str(out), {"u": u}, range(8), nmv, "hex")
    blocks = N.read_fld(str(out))
    X, ids, c, nm, nq = N.expansion_arrays(m, blocks, "u")
    assert list(nm) == list(nmv) and list(nq) == list(nmv + 1)
    assert np.array_equal(c, np.concatenate(u))
    from ithaca_hip import assembly
    l2g, sign, bnd, ng = assembly.assembly_map(ids, nm, "hex")
    assert len(l2g) == len(c)


def test_variable_order_session(tmp_path):
    """A session with two expansions on two composites (NUMMODES 4 on elements 0-3, 6 on
    4-7): the per-element orders feed the variable-order numbering, whose random global
    fields are continuous across the element faces in the session's own local frames."""
    text = open(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml")).read()
    text = text.replace('<C ID="0"> H[0-7] </C>', '<C ID="0"> H[0-3] </C>\n<C ID="7"> H[4-7] </C>')
    text = text.replace('<E COMPOSITE="C[0]" NUMMODES="4" FIELDS="u,v,w,p" TYPE="MODIFIED" />',
                        '<E COMPOSITE="C[0]" NUMMODES="4" FIELDS="u,v,w,p" TYPE="MODIFIED" />\n'
                        '<E COMPOSITE="C[7]" NUMMODES="6" FIELDS="u,v,w,p" TYPE="MODIFIED" />')
    p = tmp_path / "varp.xml"
    p.write_text(text)
    m = N.read_mesh(str(p))
    nmv = m.element_orders(field="u")
    assert list(nmv) == [4, 4, 4, 4, 6, 6, 6, 6]
    from test_assembly_varp import _continuity
    assert _continuity("hex", m.element_coords(), m.vertex_ids(), nmv, 8) < 1e-12
    with pytest.raises(N.NektarIOError, match="no expansion"):
        N.read_mesh(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml")).element_orders([99])


# The printed norms of utilities/FieldConvert/Tests/Hex_channel_C0helmsmoothing.tst:
# (field, L2, tolerance, "L inf", tolerance).
SMOOTHING_GOLDEN = [("x", 0.57735, 1e-4, 1.0, 1e-4), ("y", 0.57735, 1e-4, 1.0, 1e-4),
                    ("z", 0.57735, 1e-4, 1.0, 1e-4), ("u", 0.635792, 1e-6, 0.739606, 1e-6),
                    ("v", 0.635792, 1e-6, 0.739606, 1e-6), ("w", 0.635792, 1e-6, 0.739606, 1e-6),
                    ("p", 1.10303, 1e-6, 2.21882, 1e-6)]


def printed(x):
    """The value as the reference's output stream prints it (6 significant digits), which
    is what its tester compares with a .tst value."""
    return float(f"{x:.6g}")


def equispaced_rule(n):
    """n evenly spaced points on [-1, 1] and their weights (the exact integral of each
    Lagrange polynomial, PolyEPoints::CalculateWeights, Foundations/PolyEPoints.cpp:69-96)."""
    z = np.linspace(-1.0, 1.0, n)
    V = np.vander(z, n, increasing=True)
    mom = np.array([(1.0 - (-1.0) ** (k + 1)) / (k + 1) for k in range(n)])
    return z, np.linalg.solve(V.T, mom)


def test_helmholtz_smoothing_golden_on_the_cpu():
    """FieldConvert -m C0Projection:helmsmoothing=0.6 on the reference's Hex_channel fixture,
    restated with dense element matrices (tests/dense_ref.py) and a sparse direct solve,
    reproduces all seven L2 / "L inf" lines of Hex_channel_C0helmsmoothing.tst.  Three facts
    of the reference pipeline decide the numbers:
    * the expansion comes from the field file, whose hexahedra get nm (not nm + 1) GLL
      points per direction (MeshGraph.cpp:1179-1183; Field.hpp:141), so the smoothing runs
      on 4 points for the 4-mode field;
    * the .plt output re-expands every field on nm evenly spaced points per direction
      (OutputFileBase.cpp:83-86, MeshGraph.cpp:1463-1497) and the L2 norm integrates with
      their Newton-Cotes weights;
    * "L inf error" without an exact solution is Vmath::Vmax, the largest signed value
      (ExpList.cpp:3034-3037), not the largest magnitude.
    The smoother is ProcessC0Projection.cpp:196-224: lambda = (2 pi / 0.6)^2, forcing
    -lambda u, natural boundary conditions (a ContField of "DefaultVar" with no boundary
    expansion)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from ithaca_hip import assembly

    mesh = N.read_mesh(os.path.join(HERE, "Hex_channel_C0helmsmoothing.xml"))
    blocks = N.read_fld(os.path.join(HERE, "Hex_channel_C0helmsmoothing.fld"))
    lam = (2 * np.pi / 0.6) ** 2
    got = {}
    for f in ("u", "v", "w", "p"):
        X, ids, c, nm, nq = N.expansion_arrays(mesh, blocks, f, points="field")
        assert (nm, nq) == (4, 4)
        E, M = len(ids), nm ** 3
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
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(ng, ng))
        x = spla.spsolve(A.tocsc(), rhs)
        loc = sign * x[l2g]
        ze, we = equispaced_rule(nm)
        B1, _ = DR.modified_a(nm, ze)
        Be = np.kron(np.kron(B1, B1), B1)
        We = np.kron(np.kron(we, we), we)
        pts = [np.array((ze[i], ze[j], ze[k]))
               for k in range(nm) for j in range(nm) for i in range(nm)]
        l2 = 0.0
        vmax = -np.inf
        cmax = {a: -np.inf for a in "xyz"}
        cl2 = {a: 0.0 for a in "xyz"}
        for e in range(E):
            J = np.array([abs(np.linalg.det(DR.hex_map(X[e], p)[1])) for p in pts])
            xyz = np.array([DR.hex_map(X[e], p)[0] for p in pts])
            ue = Be @ loc[e * M:(e + 1) * M]
            l2 += np.sum(We * J * ue * ue)
            vmax = max(vmax, ue.max())
            for a, col in zip("xyz", xyz.T):
                cl2[a] += np.sum(We * J * col * col)
                cmax[a] = max(cmax[a], col.max())
        got[f] = (np.sqrt(l2), vmax)
        for a in "xyz":
            got[a] = (np.sqrt(cl2[a]), cmax[a])
    for name, l2, tl2, linf, tlinf in SMOOTHING_GOLDEN:
        print(f"{name}: L2 {got[name][0]:.7g} (golden {l2}), L inf {got[name][1]:.7g} "
              f"(golden {linf})")
        assert abs(printed(got[name][0]) - l2) <= tl2, name
        assert abs(printed(got[name][1]) - linf) <= tlinf, name
