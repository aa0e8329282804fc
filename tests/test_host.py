"""Host-side pieces of the drop-in library that need no GPU: the ear-clipping
triangulator against the vendored earcut's own index lists."""
import numpy as np

from conftest import load_golden
from solidboolean_amd import dropin


def test_earclip_matches_vendored_earcut():
    g = load_golden("earcut_vectors")
    L = dropin.load_earclip_only()
    import ctypes as C

    xy, sizes = g["xy"], g["ring_sizes"]
    pt = 0
    ring = 0
    ix = 0
    mism = []
    for case, nr in enumerate(g["rings_per_case"]):
        rs = sizes[ring:ring + nr]
        npts = int(rs.sum())
        pts = np.ascontiguousarray(xy[pt:pt + npts])
        want = g["indices"][ix:ix + g["index_counts"][case]]
        n = L.sbh_earclip(pts.reshape(-1), np.ascontiguousarray(rs, np.int64), int(nr), np.zeros(1, np.int64), 0)
        got = np.zeros(max(n, 1), np.int64)
        L.sbh_earclip(pts.reshape(-1), np.ascontiguousarray(rs, np.int64), int(nr), got, n)
        if not np.array_equal(got[:n], want):
            mism.append(case)
        pt += npts
        ring += nr
        ix += g["index_counts"][case]
    assert not mism, f"{len(mism)} polygons triangulated differently, first {mism[:5]}"


def test_earclip_simple_shapes():
    L = dropin.load_earclip_only()
    sq = np.array([[0, 0], [1, 0], [1, 1], [0, 1.0]])
    n = L.sbh_earclip(sq.reshape(-1).copy(), np.array([4], np.int64), 1, np.zeros(1, np.int64), 0)
    assert n == 6  # a square is 2 triangles
    hole = np.array([[0.4, 0.4], [0.6, 0.4], [0.6, 0.6], [0.4, 0.6]])
    xy = np.concatenate([sq * 2, hole * 2])
    n = L.sbh_earclip(xy.reshape(-1).copy(), np.array([4, 4], np.int64), 2, np.zeros(1, np.int64), 0)
    assert n == 3 * (8 + 2 - 2)  # n + 2h - 2 triangles
