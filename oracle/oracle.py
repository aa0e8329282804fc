"""ctypes binding of oracle/liboracle.so, the plain-C restatement of the
reference hot path.  Test infrastructure only: imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_P = C.c_void_p
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_libs: dict = {}


def load_lib(path: str = ORACLE_SO):
    path = os.path.abspath(path)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C oracle oracle)")
    L = C.CDLL(path)
    L.orc_normal.argtypes = [_D, _D, _D, _D]
    L.orc_prepare.argtypes = [_D, _I, C.c_size_t, _D, _D]
    L.orc_mesh_new.restype = _P
    L.orc_mesh_new.argtypes = [_D, C.c_size_t, _I, C.c_size_t]
    L.orc_mesh_free.argtypes = [_P]
    L.orc_mesh_tree_stats.argtypes = [_P, _I]
    L.orc_find_pairs.restype = C.c_size_t
    L.orc_find_pairs.argtypes = [_P, _P, _I, C.c_size_t]
    L.orc_tri_tri.restype = C.c_int
    L.orc_tri_tri.argtypes = [_D, _D, _D, _D, _D, _D, C.POINTER(C.c_int), _D, _D]
    L.orc_tri_tri_pairs.argtypes = [_P, _P, _I, C.c_size_t, _U8, _D]
    L.orc_point_in_mesh.restype = C.c_int
    L.orc_point_in_mesh.argtypes = [_P, _D, C.c_int]
    L.orc_classify_point.restype = C.c_int
    L.orc_classify_point.argtypes = [_P, _D]
    L.orc_classify_triangles.argtypes = [_P, _P, _U8, _U8]
    L.orc_position_key.argtypes = [_D, _I]
    L.orc_ray_candidate.restype = C.c_int
    L.orc_ray_candidate.argtypes = [_D, C.c_int, _D, _D, _D, _D, _D]
    L.orc_ray_candidates.restype = None
    L.orc_ray_candidates.argtypes = [C.c_size_t, _D, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), _D,
                                     np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"), _D]
    _libs[path] = L
    return L


def lib():
    return load_lib(ORACLE_SO)


def normal(a, b, c, L=None) -> np.ndarray:
    L = L or lib()
    out = np.zeros(3)
    L.orc_normal(*(np.ascontiguousarray(x, np.float64) for x in (a, b, c)), out)
    return out


def prepare(vertices: np.ndarray, triangles: np.ndarray, L=None):
    L = L or lib()
    V = np.ascontiguousarray(vertices, np.float64)
    T = np.ascontiguousarray(triangles, np.int64)
    n = T.shape[0]
    normals = np.zeros((max(n, 1), 3))
    boxes = np.zeros((max(n, 1), 9))
    L.orc_prepare(V.reshape(-1) if V.size else np.zeros(3), T.reshape(-1) if T.size else np.zeros(3, np.int64), n,
                  normals, boxes)
    return normals[:n], boxes[:n]


def tri_tri(a9, b9, L=None):
    """(returns 1 on a non-coplanar intersection, coplanar flag, source+target)"""
    L = L or lib()
    a = np.ascontiguousarray(a9, np.float64).reshape(3, 3)
    b = np.ascontiguousarray(b9, np.float64).reshape(3, 3)
    cop = C.c_int(0)
    s = np.zeros(3)
    t = np.zeros(3)
    r = L.orc_tri_tri(a[0].copy(), a[1].copy(), a[2].copy(), b[0].copy(), b[1].copy(), b[2].copy(), C.byref(cop),
                      s, t)
    return r, cop.value, np.concatenate([s, t])


def tri_tri_status(a9s, b9s, L=None):
    """intersectTwoFaces status (0 none, 1 hit, 2 coplanar) and segments for many pairs."""
    n = len(a9s)
    status = np.zeros(n, np.uint8)
    seg = np.zeros((n, 6))
    for i in range(n):
        r, cop, s = tri_tri(a9s[i], b9s[i], L)
        status[i] = 2 if cop else (1 if r else 0)
        if status[i] == 1:
            seg[i] = s
    return status, seg


def position_key(p, L=None) -> np.ndarray:
    L = L or lib()
    out = np.zeros(3, np.int64)
    L.orc_position_key(np.ascontiguousarray(p, np.float64), out)
    return out


def ray_candidate(p, axis: int, v0, v1, v2, L=None):
    """isPointInMesh's test of one candidate triangle (its Vector3::normal
    computed here): (hit, hit point)."""
    L = L or lib()
    f = lambda x: np.ascontiguousarray(x, np.float64)
    n = normal(v0, v1, v2, L)
    x = np.zeros(3)
    h = L.orc_ray_candidate(f(p), int(axis), f(v0), f(v1), f(v2), f(n), x)
    return bool(h), x


def ray_candidates(pts, axis, tris9, L=None):
    """ray_candidate over arrays: (hit flags, hit points)."""
    L = L or lib()
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    ax = np.ascontiguousarray(np.broadcast_to(np.asarray(axis, np.int32), (n,)))
    tr = np.ascontiguousarray(tris9, np.float64).reshape(n, 9)
    hit = np.zeros(max(n, 1), np.uint8)
    x = np.zeros((max(n, 1), 3))
    L.orc_ray_candidates(n, pts.reshape(-1) if n else np.zeros(3), ax, tr.reshape(-1) if n else np.zeros(9), hit,
                         x.reshape(-1))
    return hit[:n].astype(bool), x[:n]


class OracleMesh:
    def __init__(self, vertices: np.ndarray, triangles: np.ndarray, L=None):
        self._L = L or lib()
        self.vertices = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, np.int64).reshape(-1, 3)
        v = self.vertices.reshape(-1) if self.vertices.size else np.zeros(3)
        t = self.triangles.reshape(-1) if self.triangles.size else np.zeros(3, np.int64)
        self.h = self._L.orc_mesh_new(v, self.vertices.shape[0], t, self.triangles.shape[0])

    def __del__(self):
        try:
            self._L.orc_mesh_free(self.h)
        except Exception:
            pass

    @property
    def nt(self):
        return self.triangles.shape[0]

    def tree_stats(self):
        st = np.zeros(4, np.int64)
        self._L.orc_mesh_tree_stats(self.h, st)
        return dict(nodes=int(st[0]), leaves=int(st[1]), depth=int(st[2]), max_leaf=int(st[3]))

    def find_pairs(self, other: "OracleMesh") -> np.ndarray:
        n = self._L.orc_find_pairs(self.h, other.h, np.zeros(2, np.int64), 0)
        out = np.zeros((max(n, 1), 2), np.int64)
        self._L.orc_find_pairs(self.h, other.h, out, n)
        return out[:n]

    def tri_tri_pairs(self, other: "OracleMesh", pairs: np.ndarray):
        pairs = np.ascontiguousarray(pairs, np.int64).reshape(-1, 2)
        n = pairs.shape[0]
        status = np.zeros(max(n, 1), np.uint8)
        seg = np.zeros((max(n, 1), 6))
        self._L.orc_tri_tri_pairs(self.h, other.h, pairs if n else np.zeros((1, 2), np.int64), n, status, seg)
        return status[:n], seg[:n]

    def point_inside(self, p, axis: int) -> int:
        return self._L.orc_point_in_mesh(self.h, np.ascontiguousarray(p, np.float64), axis)

    def points_inside(self, pts: np.ndarray, axis: int) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        return np.array([self._L.orc_point_in_mesh(self.h, pts[i].copy(), axis) for i in range(pts.shape[0])],
                        np.uint8)

    def classify_points(self, pts: np.ndarray) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        return np.array([self._L.orc_classify_point(self.h, pts[i].copy()) for i in range(pts.shape[0])], np.uint8)

    def classify_triangles(self, other: "OracleMesh", intersected: np.ndarray | None = None) -> np.ndarray:
        out = np.zeros(max(self.nt, 1), np.uint8)
        flags = (np.zeros(max(self.nt, 1), np.uint8) if intersected is None
                 else np.ascontiguousarray(intersected, np.uint8))
        self._L.orc_classify_triangles(self.h, other.h, flags, out)
        return out[: self.nt]


def face_groups(triangles: np.ndarray, blocked: np.ndarray | None = None) -> np.ndarray:
    """Restatement of buildFaceGroups' leftover pass (src/solidboolean.cpp:
    230-238, flood fill :203-226) for triangles no loop edge touches: in index
    order, each unvisited triangle starts a group that a breadth-first search
    spreads across half-edges (edge i -> j of a triangle leads to the triangle
    that owns j -> i).  Only unblocked triangles enter the half-edge map, and
    a directed edge two of them share raises ValueError, as
    addUnintersectedTriangles (:250-285) fails on a repeated half-edge.
    Returns per triangle the index of the triangle that started its group, -1
    where blocked.  Pure Python: small meshes only."""
    tris = np.asarray(triangles, np.int64).reshape(-1, 3)
    nt = tris.shape[0]
    blk = np.zeros(nt, bool) if blocked is None else np.asarray(blocked, bool)
    owner: dict = {}
    for t in range(nt):
        if blk[t]:
            continue
        for k in range(3):
            key = (int(tris[t, k]), int(tris[t, (k + 1) % 3]))
            if key in owner:
                raise ValueError(f"repeated half-edge {key}")
            owner[key] = t
    label = np.full(nt, -1, np.int64)
    for s in range(nt):
        if blk[s] or label[s] >= 0:
            continue
        label[s] = s
        queue = [s]
        head = 0
        while head < len(queue):
            t = queue[head]
            head += 1
            for k in range(3):
                o = owner.get((int(tris[t, (k + 1) % 3]), int(tris[t, k])))
                if o is not None and label[o] < 0:
                    label[o] = s
                    queue.append(o)
    return label
