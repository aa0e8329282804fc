"""The allocation-free classification entry point against the allocating one,
through the C ABI: per-axis ray parities and votes identical on two calls in
a row (the second with fewer points), whether all three axes are asked for or
only the vote; sizes past the fixed scratch array and a target without a ray
grid are rejected."""
import ctypes as C

import numpy as np
import pytest

from solidboolean_amd import meshgen as mg
from solidboolean_amd import native as N

pytestmark = pytest.mark.gpu

LIMIT = 1 << 16


def noalloc(mesh, pts, axes: bool):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    axis = np.full((max(n, 1), 3), 7, np.uint8)
    inside = np.full(max(n, 1), 7, np.uint8)
    rc = mesh.L.sbg_classify_points_noalloc(mesh.h, pts.reshape(-1), n,
                                            axis.ctypes.data_as(C.c_void_p) if axes else None,
                                            inside.ctypes.data_as(C.c_void_p))
    return rc, axis[:n], inside[:n]


def query_points(mesh, n, seed):
    """near-surface centroids, vertices and random points in the box"""
    rng = np.random.default_rng(seed)
    tri = mesh.vertices[mesh.triangles]
    cen = tri.mean(axis=1)
    k = n // 3
    pts = np.concatenate([cen[rng.choice(len(cen), k)] * (1.0 + rng.uniform(-1e-3, 1e-3, (k, 1))),
                          mesh.vertices[rng.choice(len(mesh.vertices), k)],
                          rng.uniform(-1.6, 1.6, (n - 2 * k, 3))])
    return pts[rng.permutation(n)]


def test_noalloc_matches_allocating_entry(gpu):
    a = mg.icosphere(5)
    b = mg.transform(mg.icosphere(5), mg.random_rotation(mg.SplitMix64(4)), (0.3, 0.2, 0.1))
    A = gpu.mesh(a.vertices, a.triangles)
    B = gpu.mesh(b.vertices, b.triangles)
    big = query_points(b, 60000, 1)
    small = query_points(a, 2500, 2)
    print("allocation-free classification: 60000 then 2500 points", flush=True)
    # the allocating entry builds the ray grids
    ref_big = A.classify_points(big)
    ref_small = B.classify_points(small)
    for axes in (True, False):
        rc, axis, inside = noalloc(A, big, axes)
        assert rc == N.SBG_OK, gpu.L.sbg_last_error()
        assert np.array_equal(inside, ref_big[1])
        if axes:
            assert np.array_equal(axis, ref_big[0])
        rc, axis, inside = noalloc(B, small, axes)   # smaller, other target: nothing stale may leak through
        assert rc == N.SBG_OK, gpu.L.sbg_last_error()
        assert np.array_equal(inside, ref_small[1])
        if axes:
            assert np.array_equal(axis, ref_small[0])
    rc, _, inside = noalloc(A, big[:700], False)
    assert rc == N.SBG_OK and np.array_equal(inside, ref_big[1][:700])
    assert 0 < int(ref_big[1].sum()) < len(big)


def test_noalloc_limits(gpu):
    s = mg.icosphere(3)
    saved = gpu.tuning()
    gpu.set_tuning(early_grids=0)  # no ray grid at prepare time
    try:
        M = gpu.mesh(s.vertices, s.triangles)
    finally:
        gpu.set_tuning(**saved)
    pts = np.random.default_rng(9).uniform(-1.5, 1.5, (LIMIT + 1, 3))
    rc, _, _ = noalloc(M, pts[:10], True)
    assert rc == N.SBG_ERR_STATE   # no ray grid yet
    ref = M.classify_points(pts[:LIMIT])
    rc, _, _ = noalloc(M, pts, False)
    assert rc == N.SBG_ERR_INVALID
    rc, axis, inside = noalloc(M, pts[:LIMIT], True)
    assert rc == N.SBG_OK
    assert np.array_equal(axis, ref[0]) and np.array_equal(inside, ref[1])
    # |p| < 0.9 lies inside the unit icosphere, |p| > 1.01 outside
    r = np.linalg.norm(pts[:LIMIT], axis=1)
    assert np.all(inside[r < 0.9] == 1) and np.all(inside[r > 1.01] == 0)
    rc, _, _ = noalloc(M, pts[:0], True)
    assert rc == N.SBG_OK
