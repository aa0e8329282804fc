"""ctypes binding of oracle/_ref/libsbref.so: the reference's own sources compiled
by oracle/Makefile with the shim oracle/ref_shim.cpp.  Checker only: used to
write golden vectors (gen_golden.py) and, where the library is present, by the
tests and by bench.py's CPU-baseline leg."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

REF_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "oracle", "_ref", "libsbref.so")

_P = C.c_void_p
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def available() -> bool:
    return os.path.exists(REF_SO)


class RefLib:
    def __init__(self, path: str = REF_SO):
        L = C.CDLL(os.path.abspath(path))
        self.L = L
        L.sbref_mesh_new.restype = _P
        L.sbref_mesh_new.argtypes = [_D, C.c_size_t, _I, C.c_size_t]
        L.sbref_mesh_free.argtypes = [_P]
        L.sbref_mesh_prepare_seconds.restype = C.c_double
        L.sbref_mesh_prepare_seconds.argtypes = [_P]
        L.sbref_mesh_normals.argtypes = [_P, _D]
        L.sbref_mesh_boxes.argtypes = [_P, _D]
        L.sbref_mesh_tree_stats.argtypes = [_P, _I]
        L.sbref_find_pairs.restype = C.c_size_t
        L.sbref_find_pairs.argtypes = [_P, _P, _I, C.c_size_t, C.POINTER(C.c_double)]
        L.sbref_tri_tri.restype = C.c_int
        L.sbref_tri_tri.argtypes = [_D, _D, C.POINTER(C.c_int), _D]
        L.sbref_tri_tri_pairs.argtypes = [_P, _P, _I, C.c_size_t, _U8, _U8, _D, C.POINTER(C.c_double)]
        L.sbref_points_in_mesh.argtypes = [_P, _D, C.c_size_t, C.c_int, _U8, C.POINTER(C.c_double)]
        L.sbref_boolean_new.restype = _P
        L.sbref_boolean_new.argtypes = [_P, _P]
        L.sbref_boolean_free.argtypes = [_P]
        L.sbref_boolean_combine.restype = C.c_int
        L.sbref_boolean_combine.argtypes = [_P, C.c_int]
        L.sbref_boolean_seconds.restype = C.c_double
        L.sbref_boolean_seconds.argtypes = [_P]
        L.sbref_boolean_stage_seconds.argtypes = [_P, _D]
        L.sbref_boolean_vertices.restype = C.c_size_t
        L.sbref_boolean_vertices.argtypes = [_P, _D, C.c_size_t]
        L.sbref_boolean_fetch.restype = C.c_size_t
        L.sbref_boolean_fetch.argtypes = [_P, C.c_int, _I, C.c_size_t]
        L.sbref_boolean_intersected.restype = C.c_size_t
        L.sbref_boolean_intersected.argtypes = [_P, C.c_int, _I, C.c_size_t]
        L.sbref_boolean_groups.restype = C.c_size_t
        L.sbref_boolean_groups.argtypes = [_P, C.c_int, _I, _I, _U8, C.c_size_t]
        L.sbref_boolean_group_members.restype = C.c_size_t
        L.sbref_boolean_group_members.argtypes = [_P, C.c_int, _I, C.c_size_t]
        L.sbref_boolean_new_triangles.restype = C.c_size_t
        L.sbref_boolean_new_triangles.argtypes = [_P, _I, C.c_size_t]
        L.sbref_earcut.restype = C.c_size_t
        L.sbref_earcut.argtypes = [_D, _I, C.c_size_t, _I, C.c_size_t]
        L.sbref_load_obj.restype = C.c_int
        L.sbref_load_obj.argtypes = [C.c_char_p, _D, C.c_size_t, C.POINTER(C.c_size_t), _I,
                                     C.c_size_t, C.POINTER(C.c_size_t)]

    # -- meshes ---------------------------------------------------------------
    def mesh(self, vertices: np.ndarray, triangles: np.ndarray) -> "RefMesh":
        return RefMesh(self, vertices, triangles)

    def load_obj(self, path: str):
        nv = C.c_size_t(0)
        nt = C.c_size_t(0)
        z = np.zeros(3, np.float64)
        zi = np.zeros(3, np.int64)
        if not self.L.sbref_load_obj(path.encode(), z, 0, C.byref(nv), zi, 0, C.byref(nt)):
            raise IOError(path)
        v = np.zeros((nv.value, 3), np.float64)
        t = np.zeros((nt.value, 3), np.int64)
        self.L.sbref_load_obj(path.encode(), v, nv.value, C.byref(nv), t, nt.value, C.byref(nt))
        return v, t

    def earcut(self, rings):
        """rings: list of (k, 2) arrays, outer first.  Returns the index list."""
        xy = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64).reshape(-1, 2) for r in rings]))
        sizes = np.array([len(r) for r in rings], np.int64)
        n = self.L.sbref_earcut(xy.reshape(-1), sizes, len(rings), np.zeros(1, np.int64), 0)
        out = np.zeros(max(n, 1), np.int64)
        self.L.sbref_earcut(xy.reshape(-1), sizes, len(rings), out, n)
        return out[:n]

    def tri_tri(self, a9: np.ndarray, b9: np.ndarray):
        cop = C.c_int(0)
        seg = np.zeros(6, np.float64)
        r = self.L.sbref_tri_tri(np.ascontiguousarray(a9, np.float64).reshape(9),
                                 np.ascontiguousarray(b9, np.float64).reshape(9), C.byref(cop), seg)
        return r, cop.value, seg


class RefMesh:
    def __init__(self, lib: RefLib, vertices: np.ndarray, triangles: np.ndarray):
        self.lib = lib
        self.vertices = np.ascontiguousarray(vertices, np.float64)
        self.triangles = np.ascontiguousarray(triangles, np.int64)
        self.h = lib.L.sbref_mesh_new(self.vertices, self.vertices.shape[0],
                                      self.triangles, self.triangles.shape[0])

    def __del__(self):
        try:
            self.lib.L.sbref_mesh_free(self.h)
        except Exception:
            pass

    @property
    def prepare_seconds(self) -> float:
        return self.lib.L.sbref_mesh_prepare_seconds(self.h)

    def normals(self) -> np.ndarray:
        out = np.zeros((self.triangles.shape[0], 3), np.float64)
        self.lib.L.sbref_mesh_normals(self.h, out)
        return out

    def boxes(self) -> np.ndarray:
        out = np.zeros((self.triangles.shape[0], 9), np.float64)
        self.lib.L.sbref_mesh_boxes(self.h, out)
        return out

    def tree_stats(self):
        st = np.zeros(4, np.int64)
        self.lib.L.sbref_mesh_tree_stats(self.h, st)
        return dict(nodes=int(st[0]), leaves=int(st[1]), depth=int(st[2]), max_leaf=int(st[3]))

    def find_pairs(self, other: "RefMesh"):
        secs = C.c_double(0)
        n = self.lib.L.sbref_find_pairs(self.h, other.h, np.zeros(2, np.int64), 0, C.byref(secs))
        out = np.zeros((n, 2), np.int64)
        self.lib.L.sbref_find_pairs(self.h, other.h, out, n, C.byref(secs))
        return out, secs.value

    def tri_tri_pairs(self, other: "RefMesh", pairs: np.ndarray):
        pairs = np.ascontiguousarray(pairs, np.int64)
        n = pairs.shape[0]
        status = np.zeros(n, np.uint8)
        raw = np.zeros(n, np.uint8)
        seg = np.zeros((n, 6), np.float64)
        secs = C.c_double(0)
        self.lib.L.sbref_tri_tri_pairs(self.h, other.h, pairs, n, status, raw, seg, C.byref(secs))
        return status, raw, seg, secs.value

    def points_inside(self, pts: np.ndarray, axis: int):
        pts = np.ascontiguousarray(pts, np.float64)
        out = np.zeros(pts.shape[0], np.uint8)
        secs = C.c_double(0)
        self.lib.L.sbref_points_in_mesh(self.h, pts, pts.shape[0], axis, out, C.byref(secs))
        return out, secs.value


class RefBoolean:
    """SolidBoolean(a, b).combine() + fetch*, on the reference."""

    def __init__(self, a: RefMesh, b: RefMesh, quiet: bool = True):
        self.lib = a.lib
        self.a, self.b = a, b
        L = self.lib.L
        self.h = L.sbref_boolean_new(a.h, b.h)
        self.ok = bool(L.sbref_boolean_combine(self.h, 1 if quiet else 0))
        self.seconds = L.sbref_boolean_seconds(self.h)
        st = np.zeros(7, np.float64)
        L.sbref_boolean_stage_seconds(self.h, st)
        self.stage_seconds = st

    def __del__(self):
        try:
            self.lib.L.sbref_boolean_free(self.h)
        except Exception:
            pass

    def vertices(self) -> np.ndarray:
        L = self.lib.L
        n = L.sbref_boolean_vertices(self.h, np.zeros(3, np.float64), 0)
        out = np.zeros((n, 3), np.float64)
        L.sbref_boolean_vertices(self.h, out, n)
        return out

    def fetch(self, which: int) -> np.ndarray:
        L = self.lib.L
        n = L.sbref_boolean_fetch(self.h, which, np.zeros(3, np.int64), 0)
        out = np.zeros((n, 3), np.int64)
        L.sbref_boolean_fetch(self.h, which, out, n)
        return out

    def intersected(self, which: int) -> np.ndarray:
        L = self.lib.L
        n = L.sbref_boolean_intersected(self.h, which, np.zeros(1, np.int64), 0)
        out = np.zeros(n, np.int64)
        L.sbref_boolean_intersected(self.h, which, out, n)
        return out

    def groups(self, which: int):
        L = self.lib.L
        z = np.zeros(1, np.int64)
        n = L.sbref_boolean_groups(self.h, which, z, z, np.zeros(1, np.uint8), 0)
        sizes = np.zeros(max(n, 1), np.int64)
        firsts = np.zeros(max(n, 1), np.int64)
        sides = np.zeros(max(n, 1), np.uint8)
        L.sbref_boolean_groups(self.h, which, sizes, firsts, sides, n)
        return sizes[:n].copy(), firsts[:n].copy(), sides[:n].copy()

    def group_members(self, which: int) -> np.ndarray:
        L = self.lib.L
        n = L.sbref_boolean_group_members(self.h, which, np.zeros(1, np.int64), 0)
        out = np.zeros(max(n, 1), np.int64)
        L.sbref_boolean_group_members(self.h, which, out, n)
        return out[:n]

    def triangle_sides(self, which: int) -> np.ndarray:
        """Per ORIGINAL triangle of mesh `which`: 2 if intersected, else the side
        (0 outside / 1 inside) of the group that holds its copy.  Relies on
        addUnintersectedTriangles appending the kept triangles of mesh 0 and
        then mesh 1 in index order (solidboolean.cpp:250-286, 411-421)."""
        mesh = self.a if which == 0 else self.b
        nt = mesh.triangles.shape[0]
        inter0 = self.intersected(0)
        inter1 = self.intersected(1)
        kept0 = np.setdiff1d(np.arange(self.a.triangles.shape[0]), inter0)
        kept1 = np.setdiff1d(np.arange(self.b.triangles.shape[0]), inter1)
        start = 0 if which == 0 else len(kept0)
        kept = kept0 if which == 0 else kept1
        out = np.full(nt, 2, np.uint8)
        sizes, _, sides = self.groups(which)
        members = self.group_members(which)
        gid = np.repeat(np.arange(len(sizes)), sizes)
        sel = (members >= start) & (members < start + len(kept))
        out[kept[members[sel] - start]] = sides[gid[sel]]
        return out

    def new_triangles(self) -> np.ndarray:
        L = self.lib.L
        n = L.sbref_boolean_new_triangles(self.h, np.zeros(3, np.int64), 0)
        out = np.zeros((n, 3), np.int64)
        L.sbref_boolean_new_triangles(self.h, out, n)
        return out
