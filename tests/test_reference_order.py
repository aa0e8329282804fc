"""The host's rebuild of the reference's tree search order
(sbh_reference_pair_order): sorted into it, the candidate pairs of every
fixture that keeps them come out exactly as the compiled reference's search
reported them, and the intersecting pairs of the subdivision-6 and 1.31M
pairs likewise."""
import numpy as np
import pytest

from conftest import load_golden
from solidboolean_amd import dropin
from solidboolean_amd import meshgen as mg

FULL = ["small_s1", "small_s2", "small_s3", "case_cube-sphere", "case_simple-ring", "case_complex",
        "case_addax-and-meerkat", "cube_n8"]


def canonical(p):
    p = np.asarray(p, np.int64)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


@pytest.mark.parametrize("name", FULL)
def test_candidate_order(name):
    g = load_golden(name)
    ref = g["pairs"].astype(np.int64)
    c = canonical(ref)
    perm = dropin.reference_pair_order(g["va"], g["ta"], g["vb"], g["tb"], c)
    assert np.array_equal(c[perm], ref)


@pytest.mark.parametrize("name,seed", [("sub6_s1", 1), ("sub6_s3", 3)])
def test_hit_order_subdivision6(name, seed):
    g = load_golden(name)
    a, b = mg.icosphere_pair(6, seed, (0.5, 0.2, 0.1))
    assert mg.mesh_digest(a) == str(g["digest_a"]) and mg.mesh_digest(b) == str(g["digest_b"])
    ref = g["hit_pairs"].astype(np.int64)
    c = canonical(ref)
    perm = dropin.reference_pair_order(a.vertices, a.triangles, b.vertices, b.triangles, c)
    assert np.array_equal(c[perm], ref)


def test_hit_order_large():
    g = load_golden("large_s1")
    a, b = mg.large_pair(1)
    assert mg.mesh_digest(a) == str(g["digest_a"]) and mg.mesh_digest(b) == str(g["digest_b"])
    ref = g["hit_pairs"].astype(np.int64)
    c = canonical(ref)
    perm = dropin.reference_pair_order(a.vertices, a.triangles, b.vertices, b.triangles, c)
    assert np.array_equal(c[perm], ref)


def test_bad_index_is_an_error():
    g = load_golden("small_s1")
    with pytest.raises(ValueError, match="out of range"):
        dropin.reference_pair_order(g["va"], g["ta"], g["vb"], g["tb"], np.array([[10 ** 6, 0]]))
