"""The oracle's restatement of the reference's face grouping (leftover pass of
buildFaceGroups), pinned by closed forms: disjoint closed surfaces are one
group each, a band of blocked triangles splits a sphere into its two caps,
an open strip's groups follow its blocked cuts, and a repeated directed
edge fails as in the reference."""
import numpy as np
import pytest

from oracle.oracle import face_groups
from solidboolean_amd import meshgen as mg


def concat(meshes):
    V, T, base = [], [], 0
    for m in meshes:
        V.append(m.vertices)
        T.append(m.triangles + base)
        base += m.vertices.shape[0]
    return mg.Mesh(np.concatenate(V), np.concatenate(T))


def test_disjoint_spheres_are_one_group_each():
    parts = [mg.icosphere(2), mg.transform(mg.icosphere(1), None, (3, 0, 0)), mg.icosphere(2, 0.5)]
    m = concat(parts)
    lab = face_groups(m.triangles)
    starts = np.cumsum([0] + [p.nt for p in parts])
    for k in range(len(parts)):
        assert np.all(lab[starts[k]:starts[k + 1]] == starts[k])


def test_band_splits_sphere_into_caps():
    m = mg.icosphere(3)
    c = m.vertices[m.triangles].mean(1)
    band = np.abs(c[:, 2]) < 0.2
    lab = face_groups(m.triangles, band)
    assert np.all(lab[band] == -1)
    top, bottom = c[:, 2] >= 0.2, c[:, 2] <= -0.2
    assert len(set(lab[top])) == 1 and len(set(lab[bottom])) == 1
    assert lab[top][0] != lab[bottom][0]
    assert lab[top][0] == np.nonzero(top)[0].min() and lab[bottom][0] == np.nonzero(bottom)[0].min()


def test_strip_cut_into_pieces():
    # a strip of 2k triangles (quads split along a diagonal), consistently
    # oriented; blocking quad j cuts the strip there
    k = 10
    V = np.array([[x, y, 0.0] for x in range(k + 1) for y in (0, 1)], np.float64)
    T = []
    for x in range(k):
        a, b, c, d = 2 * x, 2 * x + 2, 2 * x + 3, 2 * x + 1
        T += [[a, b, c], [a, c, d]]
    T = np.array(T)
    assert np.all(face_groups(T) == 0)
    blocked = np.zeros(len(T), bool)
    blocked[[6, 7]] = True  # quad 3
    lab = face_groups(T, blocked)
    assert np.all(lab[:6] == 0) and np.all(lab[8:] == 8) and np.all(lab[6:8] == -1)
    # reversing triangle 9 (quad 4's second half) makes two of its directed
    # edges repeat its neighbours' ones: the reference's repeated-half-edge
    # failure; blocked, it no longer counts, and its neighbours lose it
    T2 = T.copy()
    T2[9] = T2[9][::-1]
    with pytest.raises(ValueError, match="repeated half-edge"):
        face_groups(T2)
    b9 = np.zeros(len(T), bool)
    b9[9] = True
    lab = face_groups(T2, b9)
    assert np.all(lab[:8] == 0) and lab[9] == -1 and lab[8] == 8 and np.all(lab[10:] == 8)
