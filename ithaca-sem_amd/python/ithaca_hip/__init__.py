"""Python binding of the gfx950 matrix-free operators (C ABI in include/ithaca_mf.h).

The operators mirror the reference MatrixFree layer (MatrixFreeOps/Operator.hpp):
a string key such as ``"Helmholtz_Hex_Deformed"`` selects the operator, ``set_jac`` /
``set_df`` / ``set_lambda`` configure it, and ``apply(in, out)`` is ``operator()``.
Device arrays are torch tensors on ``cuda:0`` (torch is plumbing for device memory and
streams only).  There is no CPU fallback: a missing library raises immediately.
"""
import ctypes
import os

import numpy as np
import torch  # noqa: F401  -- loads the HIP runtime before the operator library

_PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(_PKG, "..", ".."))
LIB_PATH = os.path.join(ROOT, "lib", "libithaca_mf.so")

FP64, FP32 = 0, 1
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)


class MFError(RuntimeError):
    """A non-zero status returned by the C ABI (the reference throws NekError)."""


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C ithaca-sem_amd` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i, d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.imf_create.argtypes = [ctypes.c_char_p, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                 ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                 ctypes.POINTER(_dp), i, i, ctypes.POINTER(vp)]
        L.imf_destroy.argtypes = [vp]
        L.imf_ndof.argtypes = [vp]
        L.imf_ndof.restype = d
        L.imf_needs_jac.argtypes = [vp]
        L.imf_needs_df.argtypes = [vp]
        L.imf_in_size.argtypes = [vp]
        L.imf_in_size.restype = sz
        L.imf_out_size.argtypes = [vp]
        L.imf_out_size.restype = sz
        for n in ("imf_set_jac", "imf_set_df", "imf_set_jac_device", "imf_set_df_device"):
            getattr(L, n).argtypes = [vp, vp, sz]
        L.imf_set_lambda.argtypes = [vp, d]
        L.imf_apply.argtypes = [vp, vp, vp]
        L.imf_apply_device.argtypes = [vp, vp, vp, vp]
        L.imf_apply_transpose_device.argtypes = [vp, vp, vp, vp]
        L.imf_apply_many_device.argtypes = [vp, i, vp, ctypes.c_long, vp, ctypes.c_long, vp]
        L.imf_last_error.restype = ctypes.c_char_p
        L.imf_basis_modified_a.argtypes = [i, i, vp, vp, vp, vp, vp]
        L.imf_geom_hex_device.argtypes = [vp, i, i, vp, i, vp, vp, vp]
        L.imf_geom_quad_device.argtypes = [vp, i, i, vp, i, vp, vp, vp]
        L.imf_geom_hex_aniso_device.argtypes = [vp, i, vp, vp, i, vp, vp, vp]
        L.imf_basis_modified_b.argtypes = [i, i, vp, vp, vp, vp, vp]
        L.imf_geom_tri_device.argtypes = [vp, i, i, i, i, vp, vp, vp]
        L.imf_geom_seg_device.argtypes = [vp, i, i, i, i, vp, vp, vp]
        L.imf_set_coord_dim.argtypes = [vp, i]
        L.imf_geom_prism_device.argtypes = [vp, i, i, i, i, i, vp, vp, vp]
        L.imf_geom_tet_device.argtypes = [vp, i, i, i, i, i, vp, vp, vp]
        L.imf_basis_modified_c.argtypes = [i, i, vp, vp, vp, vp, vp]
        L.imf_basis_modified_pyr_c.argtypes = [i, i, vp, vp, vp, vp, vp]
        L.imf_geom_pyr_device.argtypes = [vp, i, i, i, i, i, vp, vp, vp]
        _LIB = L
    return _LIB


def _check(status):
    if status != 0:
        raise MFError(lib().imf_last_error().decode())


def _np_ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def basis_modified_a(nm, nq):
    """Modified_A basis on nq GLL points (Basis.cpp:392-421), computed by the library."""
    z, w = np.zeros(nq), np.zeros(nq)
    D, b, db = np.zeros(nq * nq), np.zeros(nm * nq), np.zeros(nm * nq)
    _check(lib().imf_basis_modified_a(nm, nq, _np_ptr(z), _np_ptr(w), _np_ptr(D), _np_ptr(b),
                                      _np_ptr(db)))
    return dict(z=z, w=w, D=D, bdata=b, dbdata=db, nm=nm, nq=nq)


def basis_modified_b(nm, nq):
    """Modified_B basis on nq Gauss-Radau (alpha = 1, beta = 0) points (Basis.cpp:423-509),
    computed by the library: the collapsed direction of a triangle.  w are the Gauss-Radau
    weights (without the collapse factor 1/2); bdata has nm(nm+1)/2 modes."""
    M = nm * (nm + 1) // 2
    z, w = np.zeros(nq), np.zeros(nq)
    D, b, db = np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)
    _check(lib().imf_basis_modified_b(nm, nq, _np_ptr(z), _np_ptr(w), _np_ptr(D), _np_ptr(b),
                                      _np_ptr(db)))
    return dict(z=z, w=w, D=D, bdata=b, dbdata=db, nm=nm, nq=nq)


def basis_modified_c(nm, nq):
    """Modified_C basis on nq Gauss-Radau (alpha = 2, beta = 0) points (Basis.cpp:513-563),
    computed by the library: the collapsed eta2 direction of a tetrahedron.  w are the
    Gauss-Radau weights (without the collapse factor 1/4); bdata has nm(nm+1)(nm+2)/6 modes."""
    M = nm * (nm + 1) * (nm + 2) // 6
    z, w = np.zeros(nq), np.zeros(nq)
    D, b, db = np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)
    _check(lib().imf_basis_modified_c(nm, nq, _np_ptr(z), _np_ptr(w), _np_ptr(D), _np_ptr(b),
                                      _np_ptr(db)))
    return dict(z=z, w=w, D=D, bdata=b, dbdata=db, nm=nm, nq=nq)


def basis_modified_pyr_c(nm, nq):
    """ModifiedPyr_C basis on nq Gauss-Radau (alpha = 2, beta = 0) points (Basis.cpp:
    569-668), computed by the library: the collapsed eta2 direction of a pyramid.  w are the
    Gauss-Radau weights (without the collapse factor 1/4); bdata has nm(nm+1)(2nm+1)/6
    modes."""
    M = nm * (nm + 1) * (2 * nm + 1) // 6
    z, w = np.zeros(nq), np.zeros(nq)
    D, b, db = np.zeros(nq * nq), np.zeros(M * nq), np.zeros(M * nq)
    _check(lib().imf_basis_modified_pyr_c(nm, nq, _np_ptr(z), _np_ptr(w), _np_ptr(D),
                                          _np_ptr(b), _np_ptr(db)))
    return dict(z=z, w=w, D=D, bdata=b, dbdata=db, nm=nm, nq=nq)


def geom_device(shape, verts, nq, z, deformed=True, stream=None):
    """Geometric factors of straight-sided elements on the GPU (GeomFactors.cpp:399-478).

    verts: cuda float64 tensor (nElmt, 2, ncoord) for segments, (nElmt, 8, 3) for hexes,
    (nElmt, 4, 2) for quads, (nElmt, 3, 2)
    for triangles, (nElmt, 6, 3) for prisms, (nElmt, 4, 3) for tetrahedra or (nElmt, 5, 3) for
    pyramids (collapsed shapes: nq the per-direction counts).
    Returns (jac, df) cuda tensors in the GetJac / GetDerivFactors layouts."""
    verts = verts.contiguous()
    E = verts.shape[0]
    if shape == "seg":  # (nElmt, 2, ncoord) end points; nq points; z unused (linear map)
        ncoord = verts.shape[2]
        npts = E * (int(nq) if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(ncoord * npts, dtype=torch.float64, device=verts.device)
        _check(lib().imf_geom_seg_device(ctypes.c_void_p(verts.data_ptr()), E, ncoord, int(nq),
                                         int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                         ctypes.c_void_p(df.data_ptr()), _stream_ptr(stream)))
        return jac, df
    if shape == "tri":  # (nElmt, 3, 2) vertices; nq = (nq0, nq1); z unused (linear map)
        nq0, nq1 = (int(nq[0]), int(nq[1])) if isinstance(nq, (tuple, list)) else (nq, nq - 1)
        npts = E * (nq0 * nq1 if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(4 * npts, dtype=torch.float64, device=verts.device)
        _check(lib().imf_geom_tri_device(ctypes.c_void_p(verts.data_ptr()), E, nq0, nq1,
                                         int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                         ctypes.c_void_p(df.data_ptr()), _stream_ptr(stream)))
        return jac, df
    if shape == "pyr":  # (nElmt, 5, 3) vertices; nq = (nq0, nq1, nq2); z unused
        nq0, nq1, nq2 = ([int(x) for x in nq] if isinstance(nq, (tuple, list))
                         else (nq, nq, nq - 1))
        npts = E * (nq0 * nq1 * nq2 if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(9 * npts, dtype=torch.float64, device=verts.device)
        _check(lib().imf_geom_pyr_device(ctypes.c_void_p(verts.data_ptr()), E, nq0, nq1, nq2,
                                         int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                         ctypes.c_void_p(df.data_ptr()), _stream_ptr(stream)))
        return jac, df
    if shape == "tet":  # (nElmt, 4, 3) vertices; nq = (nq0, nq1, nq2); z unused (affine)
        nq0, nq1, nq2 = ([int(x) for x in nq] if isinstance(nq, (tuple, list))
                         else (nq, nq - 1, nq - 1))
        npts = E * (nq0 * nq1 * nq2 if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(9 * npts, dtype=torch.float64, device=verts.device)
        _check(lib().imf_geom_tet_device(ctypes.c_void_p(verts.data_ptr()), E, nq0, nq1, nq2,
                                         int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                         ctypes.c_void_p(df.data_ptr()), _stream_ptr(stream)))
        return jac, df
    if shape == "prism":  # (nElmt, 6, 3) vertices; nq = (nq0, nq1, nq2); z unused
        nq0, nq1, nq2 = ([int(x) for x in nq] if isinstance(nq, (tuple, list))
                         else (nq, nq, nq - 1))
        npts = E * (nq0 * nq1 * nq2 if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(9 * npts, dtype=torch.float64, device=verts.device)
        _check(lib().imf_geom_prism_device(ctypes.c_void_p(verts.data_ptr()), E, nq0, nq1, nq2,
                                           int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                           ctypes.c_void_p(df.data_ptr()), _stream_ptr(stream)))
        return jac, df
    dim = 3 if shape == "hex" else 2
    if isinstance(nq, (tuple, list)):  # hexes with nq[d] points z[d] along direction d
        if dim != 3 or len(nq) != 3 or len(z) != 3:
            raise MFError("geom_device: per-direction points need hexes, 3 counts, 3 arrays")
        npts = E * (int(np.prod(nq)) if deformed else 1)
        jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
        df = torch.empty(9 * npts, dtype=torch.float64, device=verts.device)
        zs = [np.ascontiguousarray(x, dtype=np.float64) for x in z]
        nq_a = (ctypes.c_int * 3)(*[int(n) for n in nq])
        z_a = (_dp * 3)(*[x.ctypes.data_as(_dp) for x in zs])
        _check(lib().imf_geom_hex_aniso_device(ctypes.c_void_p(verts.data_ptr()), E, nq_a, z_a,
                                               int(deformed), ctypes.c_void_p(jac.data_ptr()),
                                               ctypes.c_void_p(df.data_ptr()),
                                               _stream_ptr(stream)))
        return jac, df
    npts = E * (nq ** dim if deformed else 1)
    jac = torch.empty(npts, dtype=torch.float64, device=verts.device)
    df = torch.empty(dim * dim * npts, dtype=torch.float64, device=verts.device)
    zz = np.ascontiguousarray(z, dtype=np.float64)
    fn = lib().imf_geom_hex_device if dim == 3 else lib().imf_geom_quad_device
    _check(fn(ctypes.c_void_p(verts.data_ptr()), E, nq, _np_ptr(zz), int(deformed),
              ctypes.c_void_p(jac.data_ptr()), ctypes.c_void_p(df.data_ptr()),
              _stream_ptr(stream)))
    return jac, df


class Operator:
    """One MatrixFree operator instance: ``GetOperatorFactory().CreateInstance(key, basis,
    nElmt)`` of the reference, implemented on the GPU."""

    def __init__(self, key, nm, nq, nElmt, precision=FP64, basis=None, coord_dim=1):
        """nm, nq: the modes and points of every direction, or (hexes) tuples of the three
        directions' counts with basis a list of the three directions' tables.  Triangles
        ("*_Tri_*"): nm modes, nq = nq0 (with nq1 = nq0 - 1) or (nq0, nq1); basis the
        Modified_A and Modified_B tables of the two directions.  Prisms ("*_Prism_*"): nm
        modes, nq = nq0 (with nq1 = nq0, nq2 = nq0 - 1) or (nq0, nq1, nq2); basis the
        Modified_A, Modified_A and Modified_B tables of the three directions.  Tetrahedra
        ("*_Tet_*"): nm modes, nq = nq0 (with nq1 = nq2 = nq0 - 1) or (nq0, nq1, nq2); basis
        the Modified_A, Modified_B and Modified_C tables.  Pyramids ("*_Pyr_*"): nm modes,
        nq = nq0 (with nq1 = nq0, nq2 = nq0 - 1) or (nq0, nq1, nq2); basis the Modified_A,
        Modified_A and ModifiedPyr_C tables.  Segments ("*_Seg_*"): nm modes on nq GLL points,
        basis one Modified_A table; coord_dim (1 to 3) is the coordinate dimension of the mesh
        the segments lie in (the number of derivative factors, PhysDeriv outputs and
        IProductWRTDerivBase input  IProductWRTDerivBase_{Tri,Quad}_* take coord_dim = 3 for elements in 3D
        space (three fields, six derivative factors per point)."""
        self.key = key
        shape = key.split("_")[1]
        self.dim = 3 if shape in ("Hex", "Prism", "Tet", "Pyr") else 2
        self.nm, self.nq, self.nElmt, self.precision = nm, nq, nElmt, precision
        per_dir = isinstance(nm, (tuple, list))
        if shape == "Seg":
            # the C ABI reads direction 0's tables; dim carries the coordinate dimension
            self.dim = int(coord_dim)
            if basis is None:
                basis = basis_modified_a(nm, nq)
            nms, nqs, bases = [nm] * self.dim, [nq] * self.dim, [basis] * self.dim
        elif shape == "Tri":
            nq0, nq1 = (int(nq[0]), int(nq[1])) if isinstance(nq, (tuple, list)) else (nq, nq - 1)
            nms, nqs = [nm, nm], [nq0, nq1]
            if basis is None:
                basis = [basis_modified_a(nm, nq0), basis_modified_b(nm, nq1)]
            bases = list(basis)
        elif shape == "Pyr":
            nqs = ([int(x) for x in nq] if isinstance(nq, (tuple, list))
                   else [nq, nq, nq - 1])
            nms = [nm] * 3
            if basis is None:
                basis = [basis_modified_a(nm, nqs[0]), basis_modified_a(nm, nqs[1]),
                         basis_modified_pyr_c(nm, nqs[2])]
            bases = list(basis)
        elif shape == "Tet":
            nqs = ([int(x) for x in nq] if isinstance(nq, (tuple, list))
                   else [nq, nq - 1, nq - 1])
            nms = [nm] * 3
            if basis is None:
                basis = [basis_modified_a(nm, nqs[0]), basis_modified_b(nm, nqs[1]),
                         basis_modified_c(nm, nqs[2])]
            bases = list(basis)
        elif shape == "Prism":
            nqs = ([int(x) for x in nq] if isinstance(nq, (tuple, list)) else [nq, nq, nq - 1])
            nms = [nm] * 3
            if basis is None:
                basis = [basis_modified_a(nm, nqs[0]), basis_modified_a(nm, nqs[1]),
                         basis_modified_b(nm, nqs[2])]
            bases = list(basis)
        elif per_dir:
            nms, nqs = [int(x) for x in nm], [int(x) for x in nq]
            if basis is None:
                basis = [basis_modified_a(a, b) for a, b in zip(nms, nqs)]
            bases = list(basis)
        else:
            nms, nqs = [nm] * self.dim, [nq] * self.dim
            if basis is None:
                basis = basis_modified_a(nm, nq)
            bases = [basis] * self.dim
        self.basis = basis
        self._keep = [[np.ascontiguousarray(b[k], dtype=np.float64) for b in bases]
                      for k in ("bdata", "dbdata", "D", "w")]
        arr = lambda col: (_dp * self.dim)(*[a.ctypes.data_as(_dp) for a in col])
        nm_a = (ctypes.c_int * self.dim)(*nms)
        nq_a = (ctypes.c_int * self.dim)(*nqs)
        h = ctypes.c_void_p()
        _check(lib().imf_create(key.encode(), self.dim, nm_a, nq_a, arr(self._keep[0]),
                                arr(self._keep[1]), arr(self._keep[2]), arr(self._keep[3]),
                                nElmt, precision, ctypes.byref(h)))
        self._h = h
        if shape in ("Tri", "Quad") and int(coord_dim) == 3:  # IPDB on a surface in 3D
            _check(lib().imf_set_coord_dim(h, 3))
        self.dtype = torch.float64 if precision == FP64 else torch.float32

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _LIB is not None:
            _LIB.imf_destroy(h)
            self._h = None

    # -- reference Operator interface -------------------------------------------------
    def ndof(self):
        return lib().imf_ndof(self._h)

    def needs_jac(self):
        return bool(lib().imf_needs_jac(self._h))

    def needs_df(self):
        return bool(lib().imf_needs_df(self._h))

    @property
    def in_size(self):
        return lib().imf_in_size(self._h)

    @property
    def out_size(self):
        return lib().imf_out_size(self._h)

    def _set(self, name, a):
        if isinstance(a, torch.Tensor) and a.is_cuda:
            a = a.contiguous()
            assert a.dtype == torch.float64
            _check(getattr(lib(), f"imf_set_{name}_device")(self._h, ctypes.c_void_p(a.data_ptr()),
                                                           a.numel()))
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            _check(getattr(lib(), f"imf_set_{name}")(self._h, _np_ptr(a), a.size))

    def set_jac(self, jac):
        self._set("jac", jac)

    def set_df(self, df):
        self._set("df", df)

    def set_lambda(self, lam):
        _check(lib().imf_set_lambda(self._h, float(lam)))

    def apply(self, inp, out=None, stream=None):
        """operator()(in, out) on cuda tensors; asynchronous on the stream."""
        assert inp.is_cuda and inp.dtype == self.dtype and inp.is_contiguous()
        assert inp.numel() == self.in_size
        if out is None:
            out = torch.empty(self.out_size, dtype=self.dtype, device=inp.device)
        assert out.is_cuda and out.dtype == self.dtype and out.numel() == self.out_size
        _check(lib().imf_apply_device(self._h, ctypes.c_void_p(inp.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out

    def apply_transpose(self, inp, out=None, stream=None):
        """Euclidean transpose of the operator (adjoint), cuda tensors."""
        assert inp.is_cuda and inp.dtype == self.dtype and inp.is_contiguous()
        assert inp.numel() == self.out_size
        if out is None:
            out = torch.empty(self.in_size, dtype=self.dtype, device=inp.device)
        assert out.numel() == self.in_size
        _check(lib().imf_apply_transpose_device(self._h, ctypes.c_void_p(inp.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()),
                                                _stream_ptr(stream)))
        return out

    def apply_many(self, inp, out=None, stream=None):
        """operator()(in_v, out_v) for every row v of the 2D cuda tensor inp (rows may be
        longer than the vector: the row stride is the leading dimension)."""
        assert inp.is_cuda and inp.dtype == self.dtype and inp.dim() == 2 and inp.stride(1) == 1
        assert inp.shape[1] >= self.in_size
        if out is None:
            out = torch.empty((inp.shape[0], self.out_size), dtype=self.dtype,
                              device=inp.device)
        assert out.is_cuda and out.dtype == self.dtype and out.dim() == 2 and out.stride(1) == 1
        assert out.shape[0] == inp.shape[0] and out.shape[1] >= self.out_size
        _check(lib().imf_apply_many_device(self._h, inp.shape[0], ctypes.c_void_p(inp.data_ptr()),
                                           inp.stride(0), ctypes.c_void_p(out.data_ptr()),
                                           out.stride(0), _stream_ptr(stream)))
        return out

    def apply_host(self, inp):
        """operator()(in, out) on host numpy arrays (synchronous, FP64)."""
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        assert inp.size == self.in_size
        out = np.empty(self.out_size)
        _check(lib().imf_apply(self._h, _np_ptr(inp), _np_ptr(out)))
        return out
