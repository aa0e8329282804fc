"""Python view of the drop-in C++ API (libsolidboolean.so): SolidMesh and
SolidBoolean with the reference's calling pattern

    a = SolidMesh(vertices, triangles); a.prepare()      # GPU
    b = SolidMesh(...); b.prepare()
    sb = SolidBoolean(a, b); ok = sb.combine()
    sb.fetch_union(), sb.fetch_diff(), sb.fetch_intersect(), sb.result_vertices()
    v, t = load_obj("a.obj"); export_obj("out.obj", v, t)   # the test driver's OBJ I/O

Raises if the native libraries are missing or no GPU is usable.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import native

HOST_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsolidboolean.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_P = C.c_void_p

_lib = None

STAGES = ["searchPotentialIntersectedPairs", "processPotentialIntersectedPairs", "addUnintersectedTriangles",
          "reTriangulate", "buildPolygonsFromEdges", "buildFaceGroups", "decideGroupSide"]


def load():
    global _lib
    if _lib is not None:
        return _lib
    native.load()  # libsbgpu.so first (the host library links it)
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"{HOST_LIB} is not built")
    L = C.CDLL(HOST_LIB)
    L.sbh_last_error.restype = C.c_char_p
    L.sbh_mesh_new.restype = _P
    L.sbh_mesh_new.argtypes = [_D, C.c_size_t, _I, C.c_size_t]
    L.sbh_mesh_free.argtypes = [_P]
    L.sbh_boolean_new.restype = C.c_int
    L.sbh_boolean_new.argtypes = [_P, _P, C.POINTER(C.c_void_p)]
    L.sbh_boolean_new_opts.restype = C.c_int
    L.sbh_boolean_new_opts.argtypes = [_P, _P, C.c_uint, C.POINTER(C.c_void_p)]
    L.sbh_boolean_free.argtypes = [_P]
    L.sbh_boolean_vertices.restype = C.c_size_t
    L.sbh_boolean_vertices.argtypes = [_P, _D, C.c_size_t]
    L.sbh_boolean_fetch.restype = C.c_size_t
    L.sbh_boolean_fetch.argtypes = [_P, C.c_int, _I, C.c_size_t]
    L.sbh_boolean_triangle_sides.restype = C.c_size_t
    L.sbh_boolean_triangle_sides.argtypes = [_P, C.c_int, _U8, C.c_size_t]
    L.sbh_boolean_stage_seconds.argtypes = [_P, _D]
    L.sbh_boolean_vertex_sources.restype = C.c_size_t
    L.sbh_boolean_vertex_sources.argtypes = [_P, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                             C.c_size_t]
    _host_only_signatures(L)
    _lib = L
    return L


def earclip(rings) -> np.ndarray:
    """The host triangulator on rings (outer first) given as (k, 2) arrays."""
    L = load() if _lib is None else _lib
    xy = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64).reshape(-1, 2) for r in rings]))
    sizes = np.array([len(r) for r in rings], np.int64)
    n = L.sbh_earclip(xy.reshape(-1), sizes, len(rings), np.zeros(1, np.int64), 0)
    out = np.zeros(max(n, 1), np.int64)
    L.sbh_earclip(xy.reshape(-1), sizes, len(rings), out, n)
    return out[:n]


def load_earclip_only():
    """Loads libsolidboolean.so without touching the GPU library's device
    functions (the triangulator is pure host code)."""
    global _lib
    if _lib is None:
        L = C.CDLL(HOST_LIB)
        _host_only_signatures(L)
        return L
    return _lib


def _host_only_signatures(L):
    L.sbh_last_error.restype = C.c_char_p
    L.sbh_earclip.restype = C.c_size_t
    L.sbh_earclip.argtypes = [_D, _I, C.c_size_t, _I, C.c_size_t]
    L.sbh_obj_load.restype = C.c_int
    L.sbh_obj_load.argtypes = [C.c_char_p, _D, C.c_size_t, C.POINTER(C.c_size_t), _I, C.c_size_t,
                               C.POINTER(C.c_size_t)]
    L.sbh_obj_export.restype = C.c_int
    L.sbh_obj_export.argtypes = [C.c_char_p, _D, C.c_size_t, _I, C.c_size_t]
    _U32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    _U64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    L.sbh_reference_pair_order.restype = C.c_int
    L.sbh_reference_pair_order.argtypes = [_D, C.c_size_t, _I, C.c_size_t, _D, C.c_size_t, _I, C.c_size_t, _U32,
                                           C.c_size_t, _U64]
    L.sbh_tree_pairs.restype = C.c_size_t
    L.sbh_tree_pairs.argtypes = [_D, C.c_size_t, _I, C.c_size_t, _D, C.c_size_t, _I, C.c_size_t, _U64, C.c_size_t]


def reference_pair_order(va, ta, vb, tb, ab) -> np.ndarray:
    """The permutation that puts pairs ab (n x 2: a of mesh A, b of mesh B)
    into the order in which the reference's own tree search reports them."""
    L = load_earclip_only()
    va = np.ascontiguousarray(va, np.float64).reshape(-1, 3)
    vb = np.ascontiguousarray(vb, np.float64).reshape(-1, 3)
    ta = np.ascontiguousarray(ta, np.int64).reshape(-1, 3)
    tb = np.ascontiguousarray(tb, np.int64).reshape(-1, 3)
    ab = np.ascontiguousarray(ab, np.uint32).reshape(-1, 2)
    perm = np.zeros(max(len(ab), 1), np.uint64)
    z3, zi = np.zeros(3), np.zeros(3, np.int64)
    if not L.sbh_reference_pair_order(va.reshape(-1) if va.size else z3, len(va), ta.reshape(-1) if ta.size else zi,
                                      len(ta), vb.reshape(-1) if vb.size else z3, len(vb),
                                      tb.reshape(-1) if tb.size else zi, len(tb),
                                      ab.reshape(-1) if ab.size else np.zeros(2, np.uint32), len(ab), perm):
        raise ValueError(L.sbh_last_error().decode(errors="replace"))
    return perm[: len(ab)].astype(np.int64)


def tree_pairs(va, ta, vb, tb) -> np.ndarray:
    """Candidate pairs (n x 2) from the reference-signature tree on the host:
    AxisAlignedBoudingBoxTree's public constructor for both meshes, then
    test(root, root, boxes, pairs), in the order it reports them."""
    L = load_earclip_only()
    args = []
    for v, t in ((va, ta), (vb, tb)):
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.int64).reshape(-1, 3)
        args += [v.reshape(-1) if v.size else np.zeros(3), len(v), t.reshape(-1) if t.size else np.zeros(3, np.int64),
                 len(t)]
    n = L.sbh_tree_pairs(*args, np.zeros(2, np.uint64), 0)
    if n == 2 ** 64 - 1:
        raise ValueError(L.sbh_last_error().decode(errors="replace"))
    out = np.zeros(2 * max(n, 1), np.uint64)
    L.sbh_tree_pairs(*args, out, n)
    return out[: 2 * n].reshape(-1, 2).astype(np.int64)


def load_obj(path: str):
    """(vertices (n, 3) float64, triangles (m, 3) int64) of an OBJ file, read as
    the reference's test driver reads it (tinyobjloader, triangulating);
    raises IOError when the file cannot be read or a face index is bad."""
    L = load_earclip_only()
    nv, nt = C.c_size_t(0), C.c_size_t(0)
    if not L.sbh_obj_load(os.fsencode(path), np.zeros(3), 0, C.byref(nv), np.zeros(3, np.int64), 0, C.byref(nt)):
        raise IOError(f"{path}: {L.sbh_last_error().decode(errors='replace')}")
    v = np.zeros((max(nv.value, 1), 3))
    t = np.zeros((max(nt.value, 1), 3), np.int64)
    L.sbh_obj_load(os.fsencode(path), v.reshape(-1), nv.value, C.byref(nv), t.reshape(-1), nt.value, C.byref(nt))
    return v[: nv.value], t[: nt.value]


def export_obj(path: str, vertices, triangles) -> None:
    """Writes "v %f %f %f" and 1-based "f" lines (test/main.cpp:9-28)."""
    L = load_earclip_only()
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int64).reshape(-1, 3)
    vv = v.reshape(-1) if v.size else np.zeros(3)
    tt = t.reshape(-1) if t.size else np.zeros(3, np.int64)
    if not L.sbh_obj_export(os.fsencode(path), vv, v.shape[0], tt, t.shape[0]):
        raise IOError(f"{path}: {L.sbh_last_error().decode(errors='replace')}")


class SolidMesh:
    def __init__(self, vertices, triangles):
        self.L = load()
        self.vertices = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, np.int64).reshape(-1, 3)
        self.h = None

    def prepare(self):
        v = self.vertices.reshape(-1) if self.vertices.size else np.zeros(3)
        t = self.triangles.reshape(-1) if self.triangles.size else np.zeros(3, np.int64)
        self.h = self.L.sbh_mesh_new(v, self.vertices.shape[0], t, self.triangles.shape[0])
        if not self.h:
            raise RuntimeError(self.L.sbh_last_error().decode())

    def __del__(self):
        try:
            if self.h:
                self.L.sbh_mesh_free(self.h)
        except Exception:
            pass


class SolidBoolean:
    def __init__(self, a: SolidMesh, b: SolidMesh, reference_pair_order: bool = True, robust_weld: bool = False):
        """reference_pair_order: process the intersecting pairs in the order
        the reference's own tree search reports them, which makes the results
        the reference's bit for bit (SolidBoolean::setReferencePairOrder; on by
        default, off takes (a, b) order).  robust_weld: weld them by the
        edge-triangle crossing each one is, not by PositionKey cell
        (SolidBoolean::setRobustWeld; not the reference's behaviour)."""
        self.L = load()
        self.a, self.b = a, b
        self.h = None
        self.flags = (1 if reference_pair_order else 0) | (2 if robust_weld else 0)

    def combine(self) -> bool:
        h = C.c_void_p()
        r = self.L.sbh_boolean_new_opts(self.a.h, self.b.h, self.flags, C.byref(h))
        self.h = h
        if r < 0:
            raise RuntimeError(self.L.sbh_last_error().decode())
        return r == 1

    def __del__(self):
        try:
            if self.h:
                self.L.sbh_boolean_free(self.h)
        except Exception:
            pass

    def result_vertices(self) -> np.ndarray:
        n = self.L.sbh_boolean_vertices(self.h, np.zeros(3), 0)
        out = np.zeros((max(n, 1), 3))
        self.L.sbh_boolean_vertices(self.h, out.reshape(-1), n)
        return out[:n]

    def _fetch(self, which):
        n = self.L.sbh_boolean_fetch(self.h, which, np.zeros(3, np.int64), 0)
        out = np.zeros((max(n, 1), 3), np.int64)
        self.L.sbh_boolean_fetch(self.h, which, out.reshape(-1), n)
        return out[:n]

    def fetch_union(self):
        return self._fetch(0)

    def fetch_diff(self):
        return self._fetch(1)

    def fetch_intersect(self):
        return self._fetch(2)

    def triangle_sides(self, which: int) -> np.ndarray:
        n = self.L.sbh_boolean_triangle_sides(self.h, which, np.zeros(1, np.uint8), 0)
        out = np.zeros(max(n, 1), np.uint8)
        self.L.sbh_boolean_triangle_sides(self.h, which, out, n)
        return out[:n]

    def vertex_sources(self) -> np.ndarray:
        """(n_result_vertices, 5) indices into the input vertices (first mesh's,
        then second's): (v, ~0, ~0, ~0, ~0) for an input vertex, (x, y, p0, p1,
        p2) for a new vertex = line x-y meets plane p0-p1-p2."""
        n = self.L.sbh_boolean_vertex_sources(self.h, np.zeros(5, np.uint32), 0)
        if n == 2 ** 64 - 1:
            raise RuntimeError("a segment endpoint matched no edge-plane intersection")
        out = np.zeros((max(n, 1), 5), np.uint32)
        self.L.sbh_boolean_vertex_sources(self.h, out.reshape(-1), n)
        return out[:n]

    def stage_seconds(self) -> dict:
        out = np.zeros(7)
        self.L.sbh_boolean_stage_seconds(self.h, out)
        return dict(zip(STAGES, out.tolist()))
