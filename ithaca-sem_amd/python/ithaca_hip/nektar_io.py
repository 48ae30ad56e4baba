"""Readers for Nektar++ XML meshes and .fld field files (host side, no device work).

What is read, restated from the reference:

* the GEOMETRY section of a session/mesh file (SpatialDomains/MeshGraphXml.cpp,
  MeshGraphXmlCompressed.cpp): VERTEX, EDGE, FACE (quadrilateral faces ``Q``) and ELEMENT
  (``Q`` in 2D, ``H`` in 3D), either as text (``<V ID="i"> x y z </V>``) or compressed
  ("B64Z-LittleEndian": base64 of a zlib stream of the little-endian entity records of
  SpatialDomains/MeshEntities.hpp:54-103: int64 ids, float64 coordinates);
* the local vertex order that the reference derives from those entities: QuadGeom's
  (SpatialDomains/QuadGeom.cpp:56-86: vertex j is the end of edge j not shared with edge
  j+1) and HexGeom's (SpatialDomains/HexGeom.cpp:183-421: bottom edges from face 0, the
  vertical edges between faces 1-4, top edges from face 5, vertices chained along them);
* the EXPANSIONS section for TYPE="MODIFIED" (Modified_A with nm+1 Gauss-Lobatto-Legendre
  points per direction);
* the ELEMENTS blocks of a .fld file (LibUtilities/BasicUtils/FieldIOXml.cpp:1108-1180):
  field names, shape, basis, UNIORDER mode counts, element id ranges, and the data, one
  field after the other, each element's coefficients in its own local frame.

Written: .fld files in the reference's XML layout (FieldIOXml::v_Write,
FieldIOXml.cpp:86-330), so that fields computed here (reduced-order reconstructions,
solutions) can be post-processed with the reference's tools.

Only straight-sided quadrilateral and hexahedral meshes are read (a non-empty CURVED
section is rejected), which is the element family the operators here implement.
"""
import base64
import re
import socket
import time
import xml.etree.ElementTree as ET
import zlib
from xml.sax.saxutils import escape

import numpy as np

_VERTEX = np.dtype([("id", "<i8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
_EDGE = np.dtype([("id", "<i8"), ("v", "<i8", (2,))])
_QUAD = np.dtype([("id", "<i8"), ("e", "<i8", (4,))])
_HEX = np.dtype([("id", "<i8"), ("f", "<i8", (6,))])


class NektarIOError(ValueError):
    pass


def _b64z(text, dtype):
    raw = zlib.decompress(base64.b64decode("".join(text.split())))
    if len(raw) % dtype.itemsize:
        raise NektarIOError("compressed block size is not a multiple of its record size")
    return np.frombuffer(raw, dtype=dtype)


def _section(text, tag):
    """The <TAG ...>...</TAG> block of a file as its own XML document (the rest of a
    session file is not parsed, so unrelated sections need not be well formed)."""
    m = re.search(rf"<{tag}\b.*?</{tag}>", text, re.S)
    return ET.fromstring(m.group(0)) if m else None


def _records(parent, child, dtype, ncols):
    """Entity records of a VERTEX/EDGE/FACE/ELEMENT section.  Compressed sections hold one
    node per entity type (the section itself for VERTEX/EDGE, its Q/H children for
    FACE/ELEMENT) whose text is the B64Z stream; text sections hold one tag per entity."""
    if parent is None:
        return np.zeros(0, dtype=dtype)
    holders = [parent] if child is None else [c for c in parent if c.tag == child]
    packed = [h for h in holders if h.get("COMPRESSED")]
    if packed:
        out = []
        for h in packed:
            if h.get("COMPRESSED").upper() != "B64Z-LITTLEENDIAN" or h.get("BITSIZE", "64") != "64":
                raise NektarIOError(f"unsupported compression {h.get('COMPRESSED')!r}")
            out.append(_b64z(h.text, dtype))
        return np.concatenate(out) if out else np.zeros(0, dtype=dtype)
    tags = list(parent) if child is None else holders
    out = np.zeros(len(tags), dtype=dtype)
    names = dtype.names[1:]
    for i, el in enumerate(tags):
        vals = el.text.split()
        if len(vals) != ncols:
            raise NektarIOError(f"{el.tag} {el.get('ID')}: expected {ncols} values")
        out[i][0] = int(el.get("ID"))
        if len(names) == 1:  # one vector of ids
            out[i][1] = [int(v) for v in vals]
        else:
            for n, v in zip(names, vals):
                out[i][n] = float(v)
    return out


def _id_list(s):
    """Element id list in the reference's range syntax ("0-7,9,12-15")."""
    ids = []
    for part in s.replace(" ", "").split(","):
        if not part:
            continue
        if "-" in part:
            a, b = part.split("-")
            ids.extend(range(int(a), int(b) + 1))
        else:
            ids.append(int(part))
    return np.array(ids, dtype=np.int64)


class Mesh:
    """Straight-sided quad (2D) or hex (3D) mesh.

    elements: int64 ids in file order; element_vertex_ids[e]: the element's vertex ids in
    the reference's local order (4 or 8); coords[v]: vertex coordinates by id (dim values);
    expansions: list of dicts (composite, nummodes: an int, or a tuple of per-direction
    counts for NUMMODES="4,6,8", numpoints likewise or None, basistype / pointstype lists or
    None, type, fields)."""

    def __init__(self, dim, vertices, edges, faces, elements, composites, expansions):
        self.dim = dim
        self.coords = {int(v["id"]): np.array([v["x"], v["y"], v["z"]][:dim]) for v in vertices}
        self.edges = {int(e["id"]): tuple(int(x) for x in e["v"]) for e in edges}
        self.faces = {int(f["id"]): tuple(int(x) for x in f["e"]) for f in faces}
        self.composites = composites
        self.expansions = expansions
        if dim == 3:
            self.elements = np.array([int(h["id"]) for h in elements], dtype=np.int64)
            self.element_vertex_ids = np.array(
                [self._hex_vertices(tuple(int(x) for x in h["f"])) for h in elements],
                dtype=np.int64).reshape(-1, 8)
            self.element_face_ids = np.array([[int(x) for x in h["f"]] for h in elements],
                                             dtype=np.int64).reshape(-1, 6)
        else:
            self.elements = np.array([int(q["id"]) for q in elements], dtype=np.int64)
            self.element_vertex_ids = np.array(
                [self._quad_vertices(tuple(int(x) for x in q["e"])) for q in elements],
                dtype=np.int64).reshape(-1, 4)

    # -- local vertex order --------------------------------------------------------
    def element_orders(self, element_ids=None, field=None):
        """NUMMODES of each element (file order, or the given ids) from the EXPANSIONS
        entries and the composites they name (the reference's per-composite expansions of
        a variable-order session); field restricts to entries listing that field."""
        ids = self.elements if element_ids is None else np.asarray(element_ids, dtype=np.int64)
        order = {}
        for e in self.expansions:
            if field is not None and e["fields"] != [""] and field not in e["fields"]:
                continue
            for c in _composite_ids(e["composite"]):
                if int(c) not in self.composites:
                    raise NektarIOError(f"expansion names composite {int(c)}, not in the mesh")
                if isinstance(e["nummodes"], tuple):
                    raise NektarIOError("element_orders: the expansion has per-direction "
                                        "NUMMODES; read them from mesh.expansions")
                for el in _composite_ids(self.composites[int(c)]):
                    order[int(el)] = e["nummodes"]
        try:
            return np.array([order[int(i)] for i in ids], dtype=np.int32)
        except KeyError as exc:
            raise NektarIOError(f"element {exc.args[0]} has no expansion") from None

    def _quad_vertices(self, eids):
        verts = []
        for j in range(4):
            a, b = self.edges[eids[j]], self.edges[eids[(j + 1) % 4]]
            if a[0] in b:      # backwards: vertex 0 is shared with the next edge
                verts.append(a[1])
            elif a[1] in b:
                verts.append(a[0])
            else:
                raise NektarIOError(f"quad edges {eids[j]}, {eids[(j + 1) % 4]} share no vertex")
        return verts

    def _hex_vertices(self, fids):
        F = [self.faces[f] for f in fids]

        def shared(fa, fb):
            s = [e for e in F[fa] if e in F[fb]]
            if len(s) != 1:
                raise NektarIOError(f"hex faces {fids[fa]}, {fids[fb]} share {len(s)} edges")
            return s[0]

        E = [shared(0, f) for f in range(1, 5)]                 # bottom edges 0-3
        E += [shared(1, 4)] + [shared(f, f + 1) for f in range(1, 4)]  # vertical 4-7
        E += [shared(5, f) for f in range(1, 5)]                # top edges 8-11
        ed = [self.edges[e] for e in E]

        def first_two(a, b):
            if ed[a][0] in ed[b]:
                return [ed[a][1], ed[a][0]]
            if ed[a][1] in ed[b]:
                return [ed[a][0], ed[a][1]]
            raise NektarIOError("hex edges share no vertex")

        def chain(verts, edges):
            for i in edges:
                last = verts[-1]
                if ed[i][0] == last:
                    verts.append(ed[i][1])
                elif ed[i][1] == last:
                    verts.append(ed[i][0])
                else:
                    raise NektarIOError("hex edges share no vertex")
            return verts

        bottom = chain(first_two(0, 1), [1, 2])
        top = chain(first_two(8, 9), [9, 10])
        return bottom + top

    def boundary_faces(self, composite):
        """(element row, local face) of every face of a face composite ("F[...]" text or
        its id), local faces in the reference's hex order: 0 the xi2 = -1 face (vertices
        0-3), 1 xi1 = -1, 2 xi0 = +1, 3 xi1 = +1, 4 xi0 = -1, 5 xi2 = +1
        (SpatialDomains/HexGeom.cpp; the faces of the H records in that order)."""
        if self.dim != 3:
            raise NektarIOError("boundary_faces: hexahedral meshes only")
        text = self.composites[int(composite)] if not isinstance(composite, str) else composite
        where = {}
        for row, fids in enumerate(self.element_face_ids):
            for lf, f in enumerate(fids):
                where.setdefault(int(f), []).append((row, lf))
        out = []
        for f in _composite_ids(text):
            owners = where.get(int(f), [])
            if len(owners) != 1:
                raise NektarIOError(f"face {int(f)} belongs to {len(owners)} elements, "
                                    "not to one (not a boundary face)")
            out.append(owners[0])
        return out

    # -- helpers -------------------------------------------------------------------
    def element_coords(self, element_ids=None):
        """(E, 2^dim, dim) vertex coordinates in local order, for the given ids (default:
        file order)."""
        rows = self._rows(element_ids)
        ids = self.element_vertex_ids[rows]
        return np.array([[self.coords[int(v)] for v in r] for r in ids])

    def vertex_ids(self, element_ids=None):
        return self.element_vertex_ids[self._rows(element_ids)]

    def _rows(self, element_ids):
        if element_ids is None:
            return np.arange(len(self.elements))
        pos = {int(e): i for i, e in enumerate(self.elements)}
        try:
            return np.array([pos[int(e)] for e in element_ids])
        except KeyError as exc:
            raise NektarIOError(f"element {exc.args[0]} is not in the mesh") from None


def _composite_ids(text):
    """Ids of a composite or of an expansion's COMPOSITE attribute: every "X[...]" group of
    the text in the range syntax ("H[0-7]", "Q[1,3-5] Q[9]", "C[0,2]")."""
    return np.concatenate([_id_list(g) for g in re.findall(r"\[([^\]]*)\]", text or "")]
                          or [np.zeros(0, dtype=np.int64)])


def read_mesh(path):
    """Read the GEOMETRY and EXPANSIONS sections of a Nektar++ XML file."""
    with open(path) as fh:
        text = fh.read()
    g = _section(text, "GEOMETRY")
    if g is None:
        raise NektarIOError(f"{path}: no GEOMETRY section")
    dim = int(g.get("DIM"))
    if dim not in (2, 3):
        raise NektarIOError(f"{path}: DIM={dim} (2 or 3 supported)")
    curved = g.find("CURVED")
    if curved is not None and len(list(curved)) > 0:
        raise NektarIOError(f"{path}: curved elements are not supported")
    vertices = _records(g.find("VERTEX"), None, _VERTEX, 3)
    edges = _records(g.find("EDGE"), None, _EDGE, 2)
    el = g.find("ELEMENT")
    if dim == 3:
        faces = _records(g.find("FACE"), "Q", _QUAD, 4)
        other = [c.tag for c in (list(g.find("FACE")) if g.find("FACE") is not None else [])
                 if c.tag != "Q"]
        if other:
            raise NektarIOError(f"{path}: face types {sorted(set(other))} are not supported")
        elements = _records(el, "H", _HEX, 6)
        other = [c.tag for c in el if c.tag != "H"]
    else:
        faces = np.zeros(0, dtype=_QUAD)
        elements = _records(el, "Q", _QUAD, 4)
        other = [c.tag for c in el if c.tag != "Q"]
    if other:
        raise NektarIOError(f"{path}: element types {sorted(set(other))} are not supported")
    composites = {}
    comp = g.find("COMPOSITE")
    if comp is not None:
        for c in comp:
            composites[int(c.get("ID"))] = c.text.strip()
    expansions = []
    ex = _section(text, "EXPANSIONS")
    if ex is not None:
        def counts(v):
            # "6" -> 6; "4,6,8" (one count per direction) -> (4, 6, 8)
            vals = [int(x) for x in (v or "0").split(",") if x.strip()]
            return vals[0] if len(vals) == 1 else tuple(vals)
        for e in ex:
            expansions.append(dict(composite=e.get("COMPOSITE"), nummodes=counts(e.get("NUMMODES")),
                                   numpoints=counts(e.get("NUMPOINTS")) if e.get("NUMPOINTS") else None,
                                   basistype=(e.get("BASISTYPE") or "").split(",") if e.get("BASISTYPE") else None,
                                   pointstype=(e.get("POINTSTYPE") or "").split(",") if e.get("POINTSTYPE") else None,
                                   type=e.get("TYPE"), fields=(e.get("FIELDS") or "").split(",")))
    return Mesh(dim, vertices, edges, faces, elements, composites, expansions)


class FieldBlock:
    """One ELEMENTS block of a .fld file: fields, shape, basis, element ids, modes[e] = the
    mode counts per direction of element e (nummodes = the common ones for UNIORDER, None
    for MIXORDER), and data[f][e] = the coefficients of field f on element e (a
    (len(ids), ncoeffs) array for UNIORDER, a list of arrays for MIXORDER).  The data are
    field after field, each field element after element with nm_0 * nm_1 [* nm_2]
    coefficients per element (FieldIO::CheckFieldDefinition, FieldIO.cpp:570-760)."""

    def __init__(self, attrs, data):
        self.fields = attrs["FIELDS"].split(",")
        self.shape = attrs["SHAPE"]
        self.basis = attrs["BASIS"].split(",")
        self.ids = _id_list(attrs["ID"])
        if self.shape not in ("Hexahedron", "Quadrilateral"):
            raise NektarIOError(f"shape {self.shape!r} is not supported")
        if any(b != "Modified_A" for b in self.basis):
            raise NektarIOError(f"basis {self.basis}: only Modified_A is supported")
        spec = attrs["NUMMODESPERDIR"]
        kind, _, vals = spec.partition(":")
        vals = [int(x) for x in vals.split(",") if x.strip()]
        dim, E = len(self.basis), len(self.ids)
        if E == 0:
            raise NektarIOError("ELEMENTS block without element ids")
        if kind == "UNIORDER" and len(vals) == dim:
            self.nummodes = vals
            self.modes = np.tile(np.array(vals, dtype=np.int64), (E, 1))
        elif kind == "MIXORDER" and len(vals) == E * dim:
            self.nummodes = None
            self.modes = np.array(vals, dtype=np.int64).reshape(E, dim)
        else:
            raise NektarIOError(f"NUMMODESPERDIR {spec[:40]!r}: UNIORDER with {dim} counts or "
                                f"MIXORDER with {dim} counts per element expected")
        ncoef = self.modes.prod(axis=1)
        off = np.concatenate([[0], np.cumsum(ncoef)])
        total = int(off[-1])
        if data.size != len(self.fields) * total:
            raise NektarIOError("Input data is not the same length as header information")
        self.data = {}
        for i, f in enumerate(self.fields):
            d = data[i * total:(i + 1) * total]
            self.data[f] = (d.reshape(E, int(ncoef[0])) if self.nummodes is not None
                            else [d[off[e]:off[e + 1]] for e in range(E)])


def read_fld(path):
    """ELEMENTS blocks of a Nektar++ XML field file (.fld / .chk)."""
    with open(path) as fh:
        text = fh.read()
    blocks = []
    for m in re.finditer(r"<ELEMENTS\b([^>]*)>(.*?)</ELEMENTS>", text, re.S):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        comp = attrs.get("COMPRESSED")
        if comp:
            if comp.upper() != "B64Z-LITTLEENDIAN" or attrs.get("BITSIZE", "64") != "64":
                raise NektarIOError(f"unsupported compression {comp!r}")
            data = _b64z(m.group(2), np.dtype("<f8"))
        else:
            data = np.array(m.group(2).split(), dtype=np.float64)
        blocks.append(FieldBlock(attrs, data))
    if not blocks:
        raise NektarIOError(f"{path}: no ELEMENTS block")
    return blocks


def field_coefficients(blocks, field, element_ids):
    """Element-major coefficients (len(element_ids) * ncoeffs) of one field, gathered from
    the blocks by element id (the reference's ExpList::ExtractDataToCoeffs for a file
    whose modes match the expansion)."""
    rows = {}
    for b in blocks:
        if field not in b.data:
            continue
        for i, e in enumerate(b.ids):
            rows[int(e)] = b.data[field][i]
    try:
        return np.concatenate([rows[int(e)] for e in element_ids])
    except KeyError as exc:
        raise NektarIOError(f"field {field!r} has no data for element {exc.args[0]}") from None


def expansion_arrays(mesh, blocks, field, points="session"):
    """Inputs for the operators and the assembly of one field, in the element order of
    the file: (coords (E, 2^dim, dim) in local vertex order, vertex_ids (E, 2^dim),
    coefficients (element-major), nm, nq), with nm and nq ints when every element has the
    same order and per-element arrays otherwise (MIXORDER files of variable-order runs).
    The basis is Modified_A with Gauss-Lobatto-Legendre points, their number per direction
    set by `points`:
      "session": nm + 1, the reference's TYPE="MODIFIED" expansion of a session file;
      "field":   the expansion the reference builds from a field file's own definition
                 when it names no points (MeshGraph::SetExpansionInfo(fielddef),
                 SpatialDomains/MeshGraph.cpp:1179-1183): nm points for hexahedra and
                 nm + 1 for quadrilaterals.  FieldConvert uses it whenever a field file is
                 given without --useSessionExpansion (FieldUtils/Field.hpp:141)."""
    if points not in ("session", "field"):
        raise ValueError("points must be 'session' or 'field'")
    blk = [b for b in blocks if field in b.data]
    if not blk:
        raise NektarIOError(f"no data for field {field!r}")
    modes = np.concatenate([b.modes for b in blk])
    if np.any(modes != modes[:, :1]):
        raise NektarIOError(f"field {field!r}: elements with different mode counts per "
                            "direction are not supported")
    want = "Hexahedron" if mesh.dim == 3 else "Quadrilateral"
    if any(b.shape != want for b in blk):
        raise NektarIOError(f"field {field!r} is not on the mesh's {want} elements")
    ids = np.concatenate([b.ids for b in blk])
    nmv = modes[:, 0]
    # one order everywhere: ints; variable order: per-element arrays (Assembly takes either)
    nm = int(nmv[0]) if np.all(nmv == nmv[0]) else nmv.astype(np.int32)
    nq = nm if (points == "field" and mesh.dim == 3) else nm + 1
    return (mesh.element_coords(ids), mesh.vertex_ids(ids),
            field_coefficients(blocks, field, ids), nm, nq)


# ---------------------------------------------------------------- writing

_SHAPE_NAME = {"hex": ("Hexahedron", 3), "quad": ("Quadrilateral", 2)}


def seq_string(ids):
    """Element id list in the reference's range syntax, the inverse of _id_list
    (ParseUtils::GenerateSeqString, LibUtilities/BasicUtils/ParseUtils.h:71-106): runs of
    consecutive ids become "a-b", a repeated id is written once."""
    ids = [int(i) for i in ids]
    if not ids:
        return ""
    out = [str(ids[0])]
    first = last = ids[0]
    for i in ids:
        if i != last + 1 and i != last:
            if last != first:
                out.append(f"-{last}")
            out.append(f",{i}")
            first = i
        last = i
    if last != first:
        out.append(f"-{last}")
    return "".join(out)


def _modes_string(nmv, dim, mixed):
    """NUMMODESPERDIR value (FieldIOXml.cpp:289-330): the common counts, or every element's
    counts in element order."""
    if not mixed:
        return "UNIORDER:" + ",".join([str(int(nmv[0]))] * dim)
    return "MIXORDER:" + ",".join(str(int(m)) for m in nmv for _ in range(dim))


def b64z_encode(data):
    """"B64Z-LittleEndian" text of float64 data: a zlib stream of the little-endian values
    (default compression level), padded with zero bytes to a multiple of 3, in base64
    (CompressData::ZlibEncodeToBase64Str and BinaryStrToBase64Str,
    LibUtilities/BasicUtils/CompressData.cpp:108-133)."""
    comp = zlib.compress(np.ascontiguousarray(data, dtype="<f8").tobytes())
    comp += b"\0" * (-len(comp) % 3)
    return base64.b64encode(comp).decode("ascii")


def write_fld(path, fields, element_ids, nm, shape="hex", metadata=None, compressed=True):
    """Write one ELEMENTS block in the reference's .fld XML layout (FieldIOXml::v_Write,
    FieldIOXml.cpp:86-330, with FieldIO::AddInfoTag's Metadata section, FieldIO.cpp:
    348-393): fields = {name: element-major coefficients in each element's local frame (an
    array, or a list of per-element arrays)}, in the file's field order; modified basis
    with nm modes per direction: an int (UNIORDER) or one order per element (MIXORDER).  metadata: extra {key: value} entries of the Metadata section
    (e.g. Time, Kinvis).  compressed=False writes the data as text, which the reference
    reads as well."""
    if shape not in _SHAPE_NAME:
        raise NektarIOError(f"shape {shape!r}: hex or quad")
    sname, dim = _SHAPE_NAME[shape]
    ids = [int(i) for i in element_ids]
    if not ids or not fields:
        raise NektarIOError("write_fld: no elements or no fields")
    if not all(re.fullmatch(r"[A-Za-z_][\w]*", k) for k in list(fields) + list(metadata or {})):
        raise NektarIOError("write_fld: field and metadata names must be XML names")
    mixed = np.ndim(nm) != 0
    nmv = np.asarray(nm, dtype=np.int64).ravel() if mixed else np.full(len(ids), int(nm))
    if nmv.size != len(ids) or np.any(nmv < 1):
        raise NektarIOError(f"write_fld: {len(ids)} per-element orders expected")
    n = int(np.sum(nmv ** dim))
    data = []
    for name, v in fields.items():
        if isinstance(v, (list, tuple)):
            v = np.concatenate([np.ravel(x) for x in v]) if v else np.zeros(0)
        v = np.asarray(v, dtype=np.float64).ravel()
        if v.size != n:
            raise NektarIOError(f"field {name!r}: {v.size} values, {n} expected")
        data.append(v)
    data = np.concatenate(data)
    prov = {"Hostname": socket.gethostname(),
            "Timestamp": time.strftime("%d-%b-%Y %H:%M:%S")}
    lines = ['<?xml version="1.0" encoding="utf-8" ?>', "<NEKTAR>", "    <Metadata>",
             "        <Provenance>"]
    lines += [f"            <{k}>{escape(v)}</{k}>" for k, v in sorted(prov.items())]
    lines.append("        </Provenance>")
    lines += [f"        <{k}>{escape(str(v))}</{k}>" for k, v in sorted((metadata or {}).items())]
    lines.append("    </Metadata>")
    attrs = (f'FIELDS="{",".join(fields)}" SHAPE="{sname}" '
             f'BASIS="{",".join(["Modified_A"] * dim)}" '
             f'NUMMODESPERDIR="{_modes_string(nmv, dim, mixed)}" ID="{seq_string(ids)}"')
    if compressed:
        body = b64z_encode(data)
        attrs += ' COMPRESSED="B64Z-LittleEndian" BITSIZE="64"'
    else:
        body = " ".join(repr(float(x)) for x in data)
    lines.append(f"    <ELEMENTS {attrs}>{body}</ELEMENTS>")
    lines.append("</NEKTAR>")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_fld_metadata(path):
    """The Metadata section of a .fld file as {key: text} (Provenance entries under
    "Provenance/<key>")."""
    with open(path) as fh:
        text = fh.read()
    md = _section(text, "Metadata")
    out = {}
    if md is None:
        return out
    for child in md:
        if len(child):
            for g in child:
                out[f"{child.tag}/{g.tag}"] = (g.text or "").strip()
        else:
            out[child.tag] = (child.text or "").strip()
    return out
