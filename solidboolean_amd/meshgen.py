"""Deterministic synthetic closed meshes (icosphere, torus, cube grid).

Every operation is an elementwise IEEE-754 numpy operation written out term by
term (no BLAS, no fused multiply-add, no reductions whose order numpy may pick),
so the same seed gives bit-identical FP64 coordinates on every x86-64 host that
runs this numpy build.  Triangles are int64 index triples with outward,
counter-clockwise winding, the convention the boolean's ray-parity inside test
and the fetchDiff winding flip assume (reference: src/solidboolean.cpp:48-92,
529-548).

Random numbers: SplitMix64 (Steele, Lea, Flood 2014) with doubles taken as
(x >> 11) * 2**-53, and Gaussians by Box-Muller from the same stream.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream; vectorised draws advance the state."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1)."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def gaussian(self, n: int) -> np.ndarray:
        """n standard normals (Box-Muller, cosine branch, one pair per draw)."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        r = np.sqrt(-2.0 * np.log(1.0 - u1))
        return r * np.cos((2.0 * math.pi) * u2)


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float64
    triangles: np.ndarray  # (F, 3) int64

    @property
    def nv(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def nt(self) -> int:
        return int(self.triangles.shape[0])


def _normalize_rows(v: np.ndarray) -> np.ndarray:
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    length = np.sqrt(x * x + y * y + z * z)
    return np.stack([x / length, y / length, z / length], axis=1)


_ICO_T = (1.0 + math.sqrt(5.0)) / 2.0
_ICO_V = np.array(
    [
        [-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
        [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
        [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1],
    ],
    dtype=np.float64,
)
_ICO_F = np.array(
    [
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ],
    dtype=np.int64,
)


def icosphere(level: int, radius: float = 1.0) -> Mesh:
    """Unit icosahedron refined `level` times by 4:1 midpoint subdivision with
    midpoints projected to the sphere; V = 10*4**level + 2, F = 20*4**level.
    New vertices are numbered by ascending (min, max) edge key; the 4 children
    of face f are faces 4f..4f+3."""
    v = _normalize_rows(_ICO_V)
    f = _ICO_F.copy()
    for _ in range(level):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1)
        lo = np.minimum(e[..., 0], e[..., 1])
        hi = np.maximum(e[..., 0], e[..., 1])
        key = lo * np.int64(nv) + hi
        uniq, inv = np.unique(key.reshape(-1), return_inverse=True)
        ulo = uniq // nv
        uhi = uniq % nv
        mid = (v[ulo] + v[uhi]) * 0.5
        v = np.concatenate([v, _normalize_rows(mid)], 0)
        m = inv.reshape(-1, 3) + nv  # ab, bc, ca
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.stack(
            [
                np.stack([a, ab, ca], 1),
                np.stack([b, bc, ab], 1),
                np.stack([c, ca, bc], 1),
                np.stack([ab, bc, ca], 1),
            ],
            1,
        ).reshape(-1, 3)
    return Mesh(v * radius, f)


def torus(major: float, minor: float, nu: int, nv: int) -> Mesh:
    """Torus around z: nu segments around the axis, nv around the tube, each quad
    split into 2 triangles; V = nu*nv, F = 2*nu*nv."""
    i = np.arange(nu, dtype=np.float64)
    j = np.arange(nv, dtype=np.float64)
    theta = (2.0 * math.pi / nu) * i
    phi = (2.0 * math.pi / nv) * j
    ct, st = np.cos(theta)[:, None], np.sin(theta)[:, None]
    cp, sp = np.cos(phi)[None, :], np.sin(phi)[None, :]
    ring = major + minor * cp
    x = ring * ct
    y = ring * st
    z = np.broadcast_to(minor * sp, (nu, nv))
    verts = np.stack([x, y, z], -1).reshape(-1, 3).copy()
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i1 = (ii + 1) % nu
    j1 = (jj + 1) % nv
    v00 = ii * nv + jj
    v10 = i1 * nv + jj
    v11 = i1 * nv + j1
    v01 = ii * nv + j1
    t0 = np.stack([v00, v10, v11], -1).reshape(-1, 3)
    t1 = np.stack([v00, v11, v01], -1).reshape(-1, 3)
    tris = np.stack([t0, t1], 1).reshape(-1, 3).astype(np.int64)
    return Mesh(verts, tris)


def cube_grid(n: int, size: float = 1.0) -> Mesh:
    """Axis-aligned cube [0, size]^3 with each face an n x n grid of quads (2
    triangles each) sharing edge and corner vertices: V = 6n^2 + 2, F = 12n^2.
    Coordinates are k*size/n, exact when size/n is a power of two."""
    faces = []
    g = np.arange(n + 1, dtype=np.int64)
    uu, vv = np.meshgrid(g, g, indexing="ij")
    keys_all = []
    # (axis, fixed value, u-axis, v-axis) with u x v pointing outward
    specs = [
        (0, n, 1, 2), (0, 0, 2, 1),
        (1, n, 2, 0), (1, 0, 0, 2),
        (2, n, 0, 1), (2, 0, 1, 0),
    ]
    quads = []
    for axis, fixed, ua, va in specs:
        ijk = np.zeros((n + 1, n + 1, 3), dtype=np.int64)
        ijk[..., axis] = fixed
        ijk[..., ua] = uu
        ijk[..., va] = vv
        keys = (ijk[..., 0] * (n + 1) + ijk[..., 1]) * (n + 1) + ijk[..., 2]
        keys_all.append(keys.reshape(-1))
        quads.append(keys)
    uniq = np.unique(np.concatenate(keys_all))
    k2 = uniq // ((n + 1) * (n + 1))
    k1 = (uniq // (n + 1)) % (n + 1)
    k0 = uniq % (n + 1)
    step = size / n
    verts = np.stack([k2 * step, k1 * step, k0 * step], 1).astype(np.float64)
    for keys in quads:
        idx = np.searchsorted(uniq, keys)
        a = idx[:-1, :-1]
        b = idx[1:, :-1]
        c = idx[1:, 1:]
        d = idx[:-1, 1:]
        t0 = np.stack([a, b, c], -1).reshape(-1, 3)
        t1 = np.stack([a, c, d], -1).reshape(-1, 3)
        faces.append(np.stack([t0, t1], 1).reshape(-1, 3))
    return Mesh(verts, np.concatenate(faces, 0).astype(np.int64))


def radial_noise(mesh: Mesh, amplitude: float, rng: SplitMix64) -> Mesh:
    """v <- v * (1 + amplitude * U(-1, 1)), one draw per vertex."""
    u = rng.uniform(mesh.nv)
    s = 1.0 + amplitude * (2.0 * u - 1.0)
    return Mesh(mesh.vertices * s[:, None], mesh.triangles)


def random_rotation(rng: SplitMix64) -> np.ndarray:
    """3x3 rotation from a unit quaternion of 4 Gaussians (uniform on SO(3))."""
    g = rng.gaussian(4)
    w, x, y, z = (float(t) for t in g)
    n = math.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ],
        dtype=np.float64,
    )


def transform(mesh: Mesh, rot: np.ndarray | None, offset) -> Mesh:
    """v <- R v + t, each row written as ((R0 x + R1 y) + R2 z) + t."""
    v = mesh.vertices
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    if rot is not None:
        r = rot
        nx = r[0, 0] * x + r[0, 1] * y + r[0, 2] * z
        ny = r[1, 0] * x + r[1, 1] * y + r[1, 2] * z
        nz = r[2, 0] * x + r[2, 1] * y + r[2, 2] * z
        x, y, z = nx, ny, nz
    t = [float(c) for c in offset]
    return Mesh(np.stack([x + t[0], y + t[1], z + t[2]], 1), mesh.triangles)


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def icosphere_pair(level: int, seed: int, offset, noise: float = 0.0,
                   radius_a: float = 1.0, radius_b: float = 1.0):
    """Sphere A at the origin; sphere B rotated by a seeded random rotation and
    translated by `offset`. With noise > 0, each sphere's vertices get seeded
    radial noise first (A from stream seed*2+1, B from seed*2+2)."""
    a = icosphere(level, radius_a)
    b = icosphere(level, radius_b)
    if noise > 0.0:
        a = radial_noise(a, noise, SplitMix64(seed * 2 + 1))
        b = radial_noise(b, noise, SplitMix64(seed * 2 + 2))
    rot = random_rotation(SplitMix64(seed))
    b = transform(b, rot, offset)
    return a, b


def small_pair(seed: int = 1):
    """Two subdivision-3 unit icospheres (1,280 triangles each), B rotated and
    offset by (0.5, 0.3, 0.2)."""
    return icosphere_pair(3, seed, (0.5, 0.3, 0.2))


def large_pair(seed: int = 1, level: int = 8):
    """Two subdivision-8 unit icospheres (1,310,720 triangles each), radial noise
    1e-3 on both, B rotated and offset by (0.4, 0.25, 0.1)."""
    return icosphere_pair(level, seed, (0.4, 0.25, 0.1), noise=1e-3)


def torus_sphere_pair(seed: int = 1, nu: int = 2048, nv: int = 1024, level: int = 7):
    """Torus R=1, r=0.3 (nu x nv quads, 2 triangles each) against a radius-0.8
    icosphere of subdivision `level` centred at (1, 0, 0) with 1e-4 radial noise."""
    a = torus(1.0, 0.3, nu, nv)
    b = radial_noise(icosphere(level, 0.8), 1e-4, SplitMix64(seed * 2 + 2))
    b = transform(b, None, (1.0, 0.0, 0.0))
    return a, b


def cube_pair(n: int = 512):
    """Unit cube with n x n grid faces against the same cube shifted by
    (1e-6, 0.5, 0.5)."""
    a = cube_grid(n)
    b = transform(cube_grid(n), None, (1e-6, 0.5, 0.5))
    return a, b


BATCH_LEVELS = (1, 2, 3, 4, 5, 6)


def batch_pairs(npairs: int = 4096, seed: int = 1, levels=BATCH_LEVELS):
    """Independent icosphere pairs of uneven sizes, all draws from one
    SplitMix64(seed) stream, per pair in this order: the subdivision level
    (P(level) proportional to 4^-level over `levels`), radius A and radius B
    (uniform in [0.5, 2]), B's offset (uniform in [-0.8, 0.8]^3 times radius A),
    then B's rotation (4 Gaussians).  Both spheres of a pair share the level."""
    rng = SplitMix64(seed)
    lv = np.asarray(levels, np.int64)
    p = 4.0 ** (-lv.astype(np.float64))
    cdf = np.cumsum(p / p.sum())
    base = {int(l): icosphere(int(l)) for l in lv}
    out = []
    for _ in range(npairs):
        u = float(rng.uniform(1)[0])
        level = int(lv[min(int(np.searchsorted(cdf, u, side="right")), len(lv) - 1)])
        ra, rb = (0.5 + 1.5 * rng.uniform(2)).tolist()
        off = ((2.0 * rng.uniform(3) - 1.0) * 0.8 * ra).tolist()
        rot = random_rotation(rng)
        s = base[level]
        a = Mesh(s.vertices * ra, s.triangles)
        b = transform(Mesh(s.vertices * rb, s.triangles), rot, off)
        out.append((a, b))
    return out


def mesh_digest(mesh: Mesh) -> str:
    import hashlib
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(mesh.vertices, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(mesh.triangles, dtype="<i8").tobytes())
    return h.hexdigest()
