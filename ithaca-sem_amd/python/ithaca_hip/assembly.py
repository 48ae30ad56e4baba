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
        L.imf_assembly_nglobal.argtypes = [vp]
        L.imf_assembly_nglobal.restype = l_
        L.imf_assembly_destroy.argtypes = [vp]
        L.imf_global_to_local_device.argtypes = [vp, vp, vp, vp]
        L.imf_assemble_device.argtypes = [vp, vp, vp, vp]
        L.imf_helmsolve_cg_device.argtypes = [vp, vp, vp, vp, d, i, ctypes.POINTER(i),
                                              ctypes.POINTER(d), vp]
        L.imf_assembly_last_error.restype = ctypes.c_char_p
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


def assembly_map(vertex_ids, nm, shape):
    """(l2g, sign, on_boundary, nglobal) of the C0 numbering (host only, no device)."""
    ids = np.ascontiguousarray(vertex_ids, dtype=np.int64)
    if ids.ndim != 2 or ids.shape[1] != _nverts(shape):
        raise MFError(f"assembly: vertex ids must be (nElmt, {_nverts(shape)}) for {shape}")
    E = ids.shape[0]
    n = E * nm ** (3 if shape == "hex" else 2)
    l2g = np.empty(n, dtype=np.int32)
    sign = np.empty(n, dtype=np.int8)
    bnd = np.empty(n, dtype=np.uint8)
    ng = ctypes.c_long()
    _check(getattr(_lib(), f"imf_assembly_map_{shape}")(
        E, nm, ids.ctypes.data_as(ctypes.c_void_p), l2g.ctypes.data_as(ctypes.c_void_p),
        sign.ctypes.data_as(ctypes.c_void_p), bnd.ctypes.data_as(ctypes.c_void_p),
        ctypes.byref(ng)))
    return l2g, sign, bnd, ng.value


def assembly_map_hex(vertex_ids, nm):
    return assembly_map(vertex_ids, nm, "hex")


def assembly_map_quad(vertex_ids, nm):
    return assembly_map(vertex_ids, nm, "quad")


class Assembly:
    """Device assembly map of a hex mesh (vertex_ids (E, 8)) or a quad mesh (shape="quad",
    vertex_ids (E, 4)); dirichlet=True fixes every mode on the boundary faces (edges) to
    zero."""

    def __init__(self, vertex_ids, nm, dirichlet=True, shape="hex"):
        ids = np.ascontiguousarray(vertex_ids, dtype=np.int64)
        if ids.ndim != 2 or ids.shape[1] != _nverts(shape):
            raise MFError(f"assembly: vertex ids must be (nElmt, {_nverts(shape)}) for {shape}")
        self.shape, self.nElmt, self.nm = shape, ids.shape[0], nm
        h = ctypes.c_void_p()
        _check(getattr(_lib(), f"imf_assembly_create_{shape}")(
            self.nElmt, nm, ids.ctypes.data_as(ctypes.c_void_p), int(dirichlet), ctypes.byref(h)))
        self._h = h
        self.nglobal = _lib().imf_assembly_nglobal(h)
        self.nlocal = self.nElmt * nm ** (3 if shape == "hex" else 2)

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

    def helmsolve(self, op, rhs, x=None, tol=1e-9, max_iter=10000, stream=None):
        """x with Assemble(H(GlobalToLocal(x))) = rhs (Dirichlet modes zero); returns
        (x, iterations, relative residual)."""
        x = torch.empty_like(rhs) if x is None else x
        it, res = ctypes.c_int(), ctypes.c_double()
        _check(_lib().imf_helmsolve_cg_device(self._h, op._h, ctypes.c_void_p(rhs.data_ptr()),
                                              ctypes.c_void_p(x.data_ptr()), float(tol),
                                              int(max_iter), ctypes.byref(it), ctypes.byref(res),
                                              _stream_ptr(stream)))
        return x, it.value, res.value
