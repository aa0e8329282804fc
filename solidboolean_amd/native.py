"""ctypes binding of libsbgpu.so (include/sbgpu.h).

There is no fallback: if the HIP library is not built, or no GPU is visible,
every entry point raises.  Arrays cross the boundary as plain host pointers;
device memory stays inside the library's handles.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsbgpu.so")

SBG_OK = 0
SBG_ERR_INVALID = -1
SBG_ERR_HIP = -2
SBG_ERR_OOM = -3
SBG_ERR_OVERFLOW = -4
SBG_ERR_NO_DEVICE = -5
SBG_ERR_STATE = -6

FP64 = 0
FP32_FILTERED = 1

STAGES = ["records", "sort", "bvh", "traverse", "pair_sort", "tri_tri", "compact", "classify", "grid"]

# every symbol include/sbgpu.h declares
EXPORTS = [
    "sbg_last_error", "sbg_device_count", "sbg_context_create", "sbg_context_destroy", "sbg_context_synchronize",
    "sbg_context_stream", "sbg_mesh_create", "sbg_mesh_create64", "sbg_host_alloc", "sbg_host_free", "sbg_mesh_prepare", "sbg_mesh_destroy", "sbg_mesh_counts",
    "sbg_mesh_normals", "sbg_mesh_boxes", "sbg_mesh_bvh", "sbg_find_pairs", "sbg_pairs_create", "sbg_pairs_count",
    "sbg_pairs_copy", "sbg_pairs_destroy", "sbg_tri_tri", "sbg_pairs_results", "sbg_pairs_hits",
    "sbg_tri_tri_batch", "sbg_classify_points", "sbg_classify_points_noalloc", "sbg_classify_triangles",
    "sbg_classify_pair", "sbg_components",
    "sbg_half_edges",
    "sbg_context_set_timing",
    "sbg_context_stage_ms", "sbg_bench_tri_tri", "sbg_bench_rays", "sbg_bench_rays_wave", "sbg_calibrate_copy", "sbg_ray_certify",
    "sbg_ray_hits_apart",
    "sbg_context_set_tuning", "sbg_context_get_tuning", "sbg_autotune_pair",
    "sbg_adjoint",
    "sbg_batch_create", "sbg_batch_set_vertices", "sbg_batch_set_tuning", "sbg_batch_get_tuning", "sbg_batch_run",
    "sbg_batch_last_ms", "sbg_batch_counts", "sbg_batch_candidates", "sbg_batch_sides", "sbg_batch_autotune",
    "sbg_batch_set_spill", "sbg_batch_bench_tri_tri",
    "sbg_batch_destroy",
]


class SbgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sbgpu error {code}: {msg}")
        self.code = code


class BatchTuning(C.Structure):
    _fields_ = [("bfs_levels", C.c_int), ("dfs_waves", C.c_int), ("grid_density", C.c_int), ("use_graph", C.c_int)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("bfs_levels", C.c_int), ("dfs_waves", C.c_int), ("grid_density", C.c_int), ("depth_first", C.c_int),
                ("early_grids", C.c_int), ("ray_precision", C.c_int), ("ray_cell_cap", C.c_int)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class TriTriStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("tested", "exit_plane1", "exit_plane2", "exit_interval", "hits",
                                          "coplanar", "fallback_pairs", "filtered_pairs")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_SZ = C.c_size_t
_PSZ = C.POINTER(C.c_size_t)
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_U64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

_lib = None


def load(path: str = LIB_PATH):
    """Loads libsbgpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python solidboolean_amd/build.py`")
    L = C.CDLL(path)
    L.sbg_last_error.restype = C.c_char_p
    L.sbg_last_error.argtypes = []
    sig = {
        "sbg_device_count": [C.POINTER(C.c_int)],
        "sbg_context_create": [C.c_int, _PP],
        "sbg_context_destroy": [_P],
        "sbg_context_synchronize": [_P],
        "sbg_context_stream": [_P, _PP],
        "sbg_mesh_create": [_P, _D, _SZ, _U32, _SZ, _PP],
        "sbg_mesh_create64": [_P, _D, _SZ, _U64, _SZ, _PP],
        "sbg_host_alloc": [_SZ, _PP],
        "sbg_host_free": [_P],
        "sbg_mesh_prepare": [_P],
        "sbg_mesh_destroy": [_P],
        "sbg_mesh_counts": [_P, _PSZ, _PSZ],
        "sbg_mesh_normals": [_P, _D],
        "sbg_mesh_boxes": [_P, _D],
        "sbg_mesh_bvh": [_P, _P, C.POINTER(C.c_uint32), _F],
        "sbg_find_pairs": [_P, _P, _PP],
        "sbg_pairs_create": [_P, _U32, _SZ, _PP],
        "sbg_pairs_count": [_P, _PSZ],
        "sbg_pairs_copy": [_P, _U32],
        "sbg_pairs_destroy": [_P],
        "sbg_tri_tri": [_P, _P, _P, C.c_int, C.POINTER(TriTriStats)],
        "sbg_pairs_results": [_P, _P, _P],
        "sbg_pairs_hits": [_P, _PSZ, _P, _P],
        "sbg_tri_tri_batch": [_P, _D, _D, _SZ, _U8, _D],
        "sbg_classify_points": [_P, _D, _SZ, _P, _U8],
        "sbg_classify_points_noalloc": [_P, _D, _SZ, _P, _P],
        "sbg_classify_triangles": [_P, _P, _P, C.c_int, _U8],
        "sbg_classify_pair": [_P, _P, _P, _U8, _U8],
        "sbg_components": [_P, _P, _U32],
        "sbg_half_edges": [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, C.POINTER(C.c_uint32)],
        "sbg_context_set_timing": [_P, C.c_int],
        "sbg_context_stage_ms": [_P, _D, C.c_int],
        "sbg_bench_tri_tri": [_P, _P, _P, C.c_int, C.c_int, C.POINTER(C.c_double)],
        "sbg_bench_rays": [_P, _P, _P, C.c_int, C.c_int, C.POINTER(C.c_double), _U64],
        "sbg_bench_rays_wave": [_P, _P, _P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
        "sbg_calibrate_copy": [_P, _SZ, C.c_int, C.POINTER(C.c_double)],
        "sbg_ray_certify": [_D, _P, _D, _SZ, _P, _F, _F],
        "sbg_ray_hits_apart": [_F, _F, _F, _F, _SZ, _U8],
        "sbg_context_set_tuning": [_P, C.POINTER(Tuning)],
        "sbg_context_get_tuning": [_P, C.POINTER(Tuning)],
        "sbg_autotune_pair": [_P, _P, C.c_int, C.POINTER(Tuning), C.POINTER(C.c_double)],
        "sbg_adjoint": [_P, _D, _SZ, _U32, _SZ, _U32, _SZ, C.c_double, _P, _P, _P, _P],
        "sbg_batch_create": [_P, _D, _U32, _U64, _U64, _SZ, _PP],
        "sbg_batch_set_vertices": [_P, _D],
        "sbg_batch_set_tuning": [_P, C.POINTER(BatchTuning)],
        "sbg_batch_get_tuning": [_P, C.POINTER(BatchTuning)],
        "sbg_batch_run": [_P, C.c_int],
        "sbg_batch_last_ms": [_P, C.POINTER(C.c_double)],
        "sbg_batch_counts": [_P, _PSZ, _PSZ],
        "sbg_batch_candidates": [_P, _P, _P, _P],
        "sbg_batch_sides": [_P, _U8],
        "sbg_batch_autotune": [_P, C.c_int, C.c_int, C.POINTER(BatchTuning), C.POINTER(C.c_double)],
        "sbg_batch_set_spill": [_P, C.c_uint32],
        "sbg_batch_bench_tri_tri": [_P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(TriTriStats)],
        "sbg_batch_destroy": [_P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def _check(code: int):
    if code != SBG_OK:
        raise SbgError(code, _lib.sbg_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ray_certify(pts, axis, tris9):
    """The ray kernels' FP32 certificate on the host (sbg_ray_certify):
    (verdict, t, et) per candidate."""
    L = load()
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    ax = np.ascontiguousarray(np.broadcast_to(np.asarray(axis, np.int32), (n,)))
    tr = np.ascontiguousarray(tris9, np.float64).reshape(n, 9)
    v = np.zeros(max(n, 1), np.int32)
    t = np.zeros(max(n, 1), np.float32)
    et = np.zeros(max(n, 1), np.float32)
    _check(L.sbg_ray_certify(pts.reshape(-1) if n else np.zeros(3), ax.ctypes.data_as(C.c_void_p),
                             tr.reshape(-1) if n else np.zeros(9), n, v.ctypes.data_as(C.c_void_p), t, et))
    return v[:n], t[:n], et[:n]


def ray_hits_apart(ta, ea, tb, eb):
    L = load()
    arr = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (ta, ea, tb, eb)]
    n = len(arr[0])
    out = np.zeros(max(n, 1), np.uint8)
    _check(L.sbg_ray_hits_apart(*[a if n else np.zeros(1, np.float32) for a in arr], n, out))
    return out[:n].astype(bool)


def device_count() -> int:
    L = load()
    n = C.c_int(0)
    r = L.sbg_device_count(C.byref(n))
    return n.value if r == SBG_OK else 0


class Context:
    """One device + one HIP stream."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        _check(self.L.sbg_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sbg_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(self.L.sbg_context_synchronize(self.h))

    def set_timing(self, on: bool):
        _check(self.L.sbg_context_set_timing(self.h, 1 if on else 0))

    def stage_ms(self, reset: bool = True) -> dict:
        out = np.zeros(len(STAGES), np.float64)
        _check(self.L.sbg_context_stage_ms(self.h, out, 1 if reset else 0))
        return dict(zip(STAGES, out.tolist()))

    def calibrate_copy(self, nbytes: int, iters: int) -> float:
        ms = C.c_double(0)
        _check(self.L.sbg_calibrate_copy(self.h, nbytes, iters, C.byref(ms)))
        return ms.value

    def tuning(self) -> dict:
        t = Tuning()
        _check(self.L.sbg_context_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def half_edges(self, triangles, queries=None):
        """sbg_half_edges: (opposite (n, 3) uint32, query owners, first repeat)
        of a triangle list; 0xffffffff for none."""
        t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        q = np.ascontiguousarray(np.zeros(0) if queries is None else queries, np.uint64).reshape(-1)
        opp = np.zeros((max(len(t), 1), 3), np.uint32)
        qo = np.zeros(max(len(q), 1), np.uint32)
        first = C.c_uint32(0)
        _check(self.L.sbg_half_edges(self.h, t.ctypes.data_as(C.c_void_p) if len(t) else None, len(t),
                                     q.ctypes.data_as(C.c_void_p) if len(q) else None, len(q),
                                     opp.ctypes.data_as(C.c_void_p), qo.ctypes.data_as(C.c_void_p), C.byref(first)))
        return opp[: len(t)], qo[: len(q)], first.value

    def set_tuning(self, **kw):
        t = Tuning()
        _check(self.L.sbg_context_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(self.L.sbg_context_set_tuning(self.h, C.byref(t)))

    def autotune_pair(self, a: "Mesh", b: "Mesh", reps: int = 3):
        """Coordinate search of the single-pair knobs (sbg_autotune_pair):
        times prepare both -> find_pairs -> tri_tri -> classify_pair per
        setting (re-preparing both meshes), leaves the fastest set and
        returns it with its chain time in ms."""
        t = Tuning()
        ms = C.c_double(0)
        _check(self.L.sbg_autotune_pair(a.h, b.h, reps, C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value

    def batch(self, pairs) -> "Batch":
        """pairs: sequence of (mesh A, mesh B), each mesh an object with
        `vertices` (n x 3) and `triangles` (m x 3) or a (vertices, triangles) tuple."""
        return Batch(self, pairs)

    def adjoint(self, xyz_in, src, faces, vol_weight: float = 0.0, grad_derived=None, want_grad: bool = True):
        """Topology-fixed adjoint (sbg_adjoint): returns (derived positions,
        volume, grad_in or None)."""
        xyz = np.ascontiguousarray(xyz_in, np.float64).reshape(-1, 3)
        src = np.ascontiguousarray(src, np.uint32).reshape(-1, 5)
        faces = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        nin, nd, nf = len(xyz), len(src), len(faces)
        derived = np.zeros((max(nd, 1), 3), np.float64)
        vol = np.zeros(1, np.float64)
        grad = np.zeros((max(nin, 1), 3), np.float64) if want_grad else None
        gd = None if grad_derived is None else np.ascontiguousarray(grad_derived, np.float64).reshape(nd, 3)
        _check(self.L.sbg_adjoint(self.h, xyz.reshape(-1) if nin else np.zeros(3), nin,
                                  src.reshape(-1) if nd else np.zeros(5, np.uint32), nd,
                                  faces.reshape(-1) if nf else np.zeros(3, np.uint32), nf, float(vol_weight),
                                  _ptr(gd), _ptr(derived), _ptr(vol), _ptr(grad)))
        return derived[:nd], float(vol[0]), (grad[:nin] if want_grad else None)

    def mesh(self, vertices, triangles, prepare: bool = True) -> "Mesh":
        m = Mesh(self, vertices, triangles)
        if prepare:
            m.prepare()
        return m

    def pairs(self, ab) -> "Pairs":
        ab = np.ascontiguousarray(ab, np.uint32).reshape(-1, 2)
        h = C.c_void_p()
        _check(self.L.sbg_pairs_create(self.h, ab.reshape(-1), ab.shape[0], C.byref(h)))
        return Pairs(self, h)

    def tri_tri_batch(self, a9, b9):
        a9 = np.ascontiguousarray(a9, np.float64).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float64).reshape(-1, 9)
        n = a9.shape[0]
        assert b9.shape[0] == n
        status = np.zeros(max(n, 1), np.uint8)
        seg = np.zeros((max(n, 1), 6), np.float64)
        _check(self.L.sbg_tri_tri_batch(self.h, a9.reshape(-1) if n else np.zeros(9), b9.reshape(-1) if n else np.zeros(9),
                                        n, status, seg.reshape(-1)))
        return status[:n], seg[:n]


def host_array(n: int, dtype=np.uint8) -> np.ndarray:
    """A NumPy array of n elements in page-locked host memory (sbg_host_alloc),
    freed when the array is collected.  Classification results written into
    such arrays arrive by DMA, without the staged copy other memory gets."""
    L = load()
    dt = np.dtype(dtype)
    nbytes = max(int(n), 1) * dt.itemsize
    ptr = C.c_void_p()
    _check(L.sbg_host_alloc(nbytes, C.byref(ptr)))
    buf = (C.c_uint8 * nbytes).from_address(ptr.value)
    arr = np.frombuffer(buf, dtype=dt, count=max(int(n), 1))
    weakref.finalize(buf, L.sbg_host_free, C.c_void_p(ptr.value))
    return arr


class Mesh:
    def __init__(self, ctx: Context, vertices, triangles):
        self.ctx = ctx
        self.L = ctx.L
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        t = np.asarray(triangles).reshape(-1, 3)
        if t.size and t.dtype.kind in "iu" and t.dtype.itemsize == 8:
            # 64-bit indices (SolidMesh's size_t) go as they are: the library
            # narrows them while it stages the upload and refuses any index
            # >= nverts, a negative one included (as uint64 it is huge)
            t = np.ascontiguousarray(t).view(np.uint64)
            create = self.L.sbg_mesh_create64
        else:
            if t.size and t.dtype != np.uint32 and (t.min() < 0 or t.max() >= 2 ** 32):
                raise ValueError("triangle index out of uint32 range")
            t = np.ascontiguousarray(t, np.uint32)
            create = self.L.sbg_mesh_create
        self.nv, self.nt = v.shape[0], t.shape[0]
        h = C.c_void_p()
        _check(create(ctx.h, v.reshape(-1) if v.size else np.zeros(3), self.nv,
                      t.reshape(-1) if t.size else np.zeros(3, np.uint32), self.nt, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sbg_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self):
        _check(self.L.sbg_mesh_prepare(self.h))

    def normals(self) -> np.ndarray:
        out = np.zeros((max(self.nt, 1), 3), np.float64)
        _check(self.L.sbg_mesh_normals(self.h, out.reshape(-1)))
        return out[: self.nt]

    def boxes(self) -> np.ndarray:
        out = np.zeros((max(self.nt, 1), 6), np.float64)
        _check(self.L.sbg_mesh_boxes(self.h, out.reshape(-1)))
        return out[: self.nt]

    def bvh(self):
        """(nodes as a structured array, root ref, root box)"""
        dt = np.dtype([("box0", np.float32, 6), ("box1", np.float32, 6), ("c0", np.uint32), ("c1", np.uint32),
                       ("parent", np.int32), ("pad", np.int32)])
        nodes = np.zeros(max(self.nt - 1, 1), dt)
        root = C.c_uint32(0)
        rb = np.zeros(6, np.float32)
        _check(self.L.sbg_mesh_bvh(self.h, nodes.ctypes.data_as(C.c_void_p), C.byref(root), rb))
        return nodes[: max(self.nt - 1, 0)], root.value, rb

    def find_pairs(self, other: "Mesh") -> "Pairs":
        h = C.c_void_p()
        _check(self.L.sbg_find_pairs(self.h, other.h, C.byref(h)))
        return Pairs(self.ctx, h)

    def classify_points(self, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        axis = np.zeros((max(n, 1), 3), np.uint8)
        inside = np.zeros(max(n, 1), np.uint8)
        _check(self.L.sbg_classify_points(self.h, pts.reshape(-1) if n else np.zeros(3), n, _ptr(axis), inside))
        return axis[:n], inside[:n]

    def classify_triangles(self, other: "Mesh", pairs: "Pairs", side: int, out=None) -> np.ndarray:
        """sbg_classify_triangles; out: optional contiguous uint8 array of at
        least max(nt, 1) entries, filled and returned."""
        if out is None:
            out = np.zeros(max(self.nt, 1), np.uint8)
        elif out.dtype != np.uint8 or not out.flags.c_contiguous or out.size < max(self.nt, 1):
            raise ValueError("out must be a contiguous uint8 array of the mesh's size")
        _check(self.L.sbg_classify_triangles(self.h, other.h, pairs.h, side, out))
        return out[: self.nt]

    def components(self, blocked=None) -> np.ndarray:
        """sbg_components: per triangle the smallest triangle index of its
        component across opposite half-edges (0xffffffff where blocked)."""
        out = np.zeros(max(self.nt, 1), np.uint32)
        bl = None if blocked is None else np.ascontiguousarray(blocked, np.uint8)
        _check(self.L.sbg_components(self.h, bl.ctypes.data_as(C.c_void_p) if bl is not None else None, out))
        return out[: self.nt]

    def classify_pair(self, other: "Mesh", pairs: "Pairs", out=None):
        """sbg_classify_pair: (votes of self's triangles, votes of other's),
        self being the pairs' first mesh.  out: optional pair of uint8
        arrays of at least max(nt, 1) entries each, filled and returned (a
        caller that classifies step after step reuses them)."""
        if out is None:
            oa = np.zeros(max(self.nt, 1), np.uint8)
            ob = np.zeros(max(other.nt, 1), np.uint8)
        else:
            oa, ob = out
            if (oa.dtype != np.uint8 or ob.dtype != np.uint8 or not oa.flags.c_contiguous
                    or not ob.flags.c_contiguous or oa.size < max(self.nt, 1) or ob.size < max(other.nt, 1)):
                raise ValueError("out must be two contiguous uint8 arrays of the meshes' sizes")
        _check(self.L.sbg_classify_pair(self.h, other.h, pairs.h, oa, ob))
        return oa[: self.nt], ob[: other.nt]

    def bench_rays(self, other: "Mesh", pairs: "Pairs", side: int, iters: int):
        """sbg_bench_rays: mean launch ms of the x/y ray kernel of
        classify_triangles, and the work of one launch."""
        ms = C.c_double(0)
        work = np.zeros(5, np.uint64)
        _check(self.L.sbg_bench_rays(self.h, other.h, pairs.h, side, iters, C.byref(ms), work))
        return ms.value, dict(zip(("rays", "scanned", "passed", "deferred", "to_fp64"), (int(x) for x in work)))

    def bench_rays_wave(self, other: "Mesh", pairs: "Pairs", side: int, iters: int):
        """sbg_bench_rays_wave: mean launch ms of the FP64 wave ray kernel over
        the rays one x/y launch defers, and their number."""
        ms = C.c_double(0)
        rays = C.c_uint64(0)
        _check(self.L.sbg_bench_rays_wave(self.h, other.h, pairs.h, side, iters, C.byref(ms), C.byref(rays)))
        return ms.value, int(rays.value)


class Pairs:
    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.L = ctx.L
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sbg_pairs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = C.c_size_t(0)
        _check(self.L.sbg_pairs_count(self.h, C.byref(n)))
        return n.value

    def array(self) -> np.ndarray:
        n = len(self)
        out = np.zeros((max(n, 1), 2), np.uint32)
        _check(self.L.sbg_pairs_copy(self.h, out.reshape(-1)))
        return out[:n]

    def tri_tri(self, a: Mesh, b: Mesh, precision: int = FP64, stats: bool = True):
        st = TriTriStats()
        _check(self.L.sbg_tri_tri(a.h, b.h, self.h, precision, C.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def results(self):
        n = len(self)
        status = np.zeros(max(n, 1), np.uint8)
        seg = np.zeros((max(n, 1), 6), np.float64)
        _check(self.L.sbg_pairs_results(self.h, _ptr(status), _ptr(seg)))
        return status[:n], seg[:n]

    def hits(self):
        nh = C.c_size_t(0)
        _check(self.L.sbg_pairs_hits(self.h, C.byref(nh), None, None))
        n = nh.value
        ab = np.zeros((max(n, 1), 2), np.uint32)
        seg = np.zeros((max(n, 1), 6), np.float64)
        _check(self.L.sbg_pairs_hits(self.h, C.byref(nh), _ptr(ab), _ptr(seg)))
        return ab[:n], seg[:n]

    def bench(self, a: Mesh, b: Mesh, precision: int, iters: int) -> float:
        ms = C.c_double(0)
        _check(self.L.sbg_bench_tri_tri(a.h, b.h, self.h, precision, iters, C.byref(ms)))
        return ms.value


def _vt(m):
    if hasattr(m, "vertices"):
        return m.vertices, m.triangles
    return m[0], m[1]


class Batch:
    """Many independent mesh pairs run as one (sbg_batch_*)."""

    def __init__(self, ctx: Context, pairs):
        self.ctx = ctx
        self.L = ctx.L
        meshes = []
        for a, b in pairs:
            meshes.append(_vt(a))
            meshes.append(_vt(b))
        vb = np.zeros(len(meshes) + 1, np.uint64)
        tb = np.zeros(len(meshes) + 1, np.uint64)
        for i, (v, t) in enumerate(meshes):
            vb[i + 1] = vb[i] + len(v)
            tb[i + 1] = tb[i] + len(t)
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64).reshape(-1, 3) for v, _ in meshes]))
        tris = np.ascontiguousarray(np.concatenate([np.asarray(t).reshape(-1, 3) for _, t in meshes]), np.uint32)
        self.vert_begin, self.tri_begin = vb, tb
        self.nmeshes = len(meshes)
        self.nt = int(tb[-1])
        h = C.c_void_p()
        _check(self.L.sbg_batch_create(ctx.h, xyz.reshape(-1), tris.reshape(-1), vb, tb, self.nmeshes, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sbg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_vertices(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        _check(self.L.sbg_batch_set_vertices(self.h, xyz))

    def tuning(self) -> dict:
        t = BatchTuning()
        _check(self.L.sbg_batch_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = BatchTuning()
        _check(self.L.sbg_batch_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(self.L.sbg_batch_set_tuning(self.h, C.byref(t)))

    def bench_tri_tri(self, precision: int, iters: int):
        """sbg_batch_bench_tri_tri: (ms per GD launch over the last run's
        candidates, statistics of one launch)."""
        ms = C.c_double(0)
        st = TriTriStats()
        _check(self.L.sbg_batch_bench_tri_tri(self.h, precision, iters, C.byref(ms), C.byref(st)))
        return ms.value, st.as_dict()

    def set_spill(self, per_wave: int):
        """Testing aid: the depth-first search's spill region per wave."""
        _check(self.L.sbg_batch_set_spill(self.h, int(per_wave)))

    def run(self, precision: int = FP32_FILTERED) -> float:
        _check(self.L.sbg_batch_run(self.h, precision))
        ms = C.c_double(0)
        _check(self.L.sbg_batch_last_ms(self.h, C.byref(ms)))
        return ms.value

    def counts(self):
        c, h = C.c_size_t(0), C.c_size_t(0)
        _check(self.L.sbg_batch_counts(self.h, C.byref(c), C.byref(h)))
        return c.value, h.value

    def candidates(self):
        n, _ = self.counts()
        ab = np.zeros((max(n, 1), 2), np.uint32)
        st = np.zeros(max(n, 1), np.uint8)
        seg = np.zeros((max(n, 1), 6), np.float64)
        _check(self.L.sbg_batch_candidates(self.h, _ptr(ab), _ptr(st), _ptr(seg)))
        return ab[:n], st[:n], seg[:n]

    def sides(self) -> np.ndarray:
        out = np.zeros(max(self.nt, 1), np.uint8)
        _check(self.L.sbg_batch_sides(self.h, out))
        return out[: self.nt]

    def autotune(self, precision: int = FP32_FILTERED, reps: int = 3):
        t = BatchTuning()
        ms = C.c_double(0)
        _check(self.L.sbg_batch_autotune(self.h, precision, reps, C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value
