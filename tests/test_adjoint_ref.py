"""The CPU adjoint checker against central differences on small synthetic
edge-plane configurations (no GPU): a closed tetrahedron-like face set whose
vertices are partly derived, and random lines against random planes."""
import numpy as np

import adjoint_ref as R


def random_config(seed, nin=12, nd=6):
    rng = np.random.default_rng(seed)
    xyz = rng.standard_normal((nin, 3))
    src = np.stack([rng.choice(nin, 5, replace=False) for _ in range(nd)]).astype(np.int64)
    faces = rng.integers(0, nin + nd, (20, 3))
    return xyz, src, faces


def test_derived_point_is_on_line_and_plane():
    xyz, src, _ = random_config(1)
    X = R.derived(xyz, src)
    x, y, p0, p1, p2 = (xyz[src[:, k]] for k in range(5))
    N = np.cross(p1 - p0, p2 - p0)
    assert np.allclose(np.einsum("ij,ij->i", X - p0, N), 0, atol=1e-12)
    assert np.allclose(np.cross(X - x, y - x), 0, atol=1e-12)


def test_volume_of_unit_tetrahedron_and_cube():
    tet = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    faces = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    assert abs(R.volume(tet, np.zeros((0, 5), np.int64), faces) - 1 / 6) < 1e-15


def test_vjp_matches_central_differences():
    for seed in range(5):
        xyz, src, faces = random_config(seed)
        W = np.random.default_rng(100 + seed).standard_normal((len(src), 3))
        g = R.vjp(xyz, src, faces, vol_weight=0.7, grad_derived=W)
        h = 1e-6
        for u in R.directions(len(xyz), 8, seed):
            phi = lambda z: 0.7 * R.volume(z, src, faces) + np.sum(W * R.derived(z, src))
            fd = (phi(xyz + h * u) - phi(xyz - h * u)) / (2 * h)
            assert abs(np.sum(g * u) - fd) <= 1e-7 * max(1.0, abs(fd)), (seed, np.sum(g * u), fd)


def test_planted_error_is_caught():
    """The same check fails for a gradient whose plane-vertex entries are
    negated, so it can tell a wrong adjoint from a right one."""
    xyz, src, faces = random_config(3)
    g = R.vjp(xyz, src, faces, vol_weight=0.0, grad_derived=np.ones((len(src), 3)))
    x, y, p0, p1, p2 = (xyz[src[:, k]] for k in range(5))
    u = R.directions(len(xyz), 1, 9)[0]
    h = 1e-6
    fd = (np.sum(R.derived(xyz + h * u, src)) - np.sum(R.derived(xyz - h * u, src))) / (2 * h)
    bad = g.copy()
    bad[src[:, 3]] = -g[src[:, 3]]
    assert abs(np.sum(g * u) - fd) < 1e-7 * max(1, abs(fd))
    assert abs(np.sum(bad * u) - fd) > 1e-5
