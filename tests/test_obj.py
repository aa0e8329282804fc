"""OBJ input and output of the reference's test driver (test/main.cpp:9-69):
the reader equals tinyobjloader as the reference compiles it (vectors written
by tests/golden/gen_obj_golden.py) on number forms, face index forms,
polygons, line endings, groups and its failures, and reads the reference's
own test cases to the arrays its combine() ran on; the writer's text reads
back to the same mesh at its 6 printed decimals."""
import os

import numpy as np
import pytest

from solidboolean_amd import dropin

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OBJ = os.path.join(GOLDEN, "obj")
CASES = ["cube-sphere", "simple-ring", "complex", "addax-and-meerkat"]

with np.load(os.path.join(GOLDEN, "obj_golden.npz"), allow_pickle=False) as _g:
    EXPECT = {k: _g[k] for k in _g.files}
NAMES = sorted({k.split("__")[0] for k in EXPECT})


@pytest.mark.parametrize("name", NAMES)
def test_reader_matches_tinyobjloader(name):
    path = os.path.join(OBJ, name + ".obj")
    if not int(EXPECT[name + "__ok"]):
        with pytest.raises(IOError, match="bad face index"):
            dropin.load_obj(path)
        return
    v, t = dropin.load_obj(path)
    ev, et = EXPECT[name + "__v"], EXPECT[name + "__t"]
    assert v.shape == ev.shape and np.array_equal(v.view(np.uint64), ev.view(np.uint64))
    assert np.array_equal(t, et.reshape(-1, 3))


@pytest.mark.parametrize("case", CASES)
def test_reference_cases_read_as_the_driver_did(case):
    with np.load(os.path.join(GOLDEN, f"case_{case}.npz"), allow_pickle=False) as g:
        for side in "ab":
            v, t = dropin.load_obj(os.path.join(OBJ, "cases", case, side + ".obj"))
            assert np.array_equal(v.view(np.uint64), g["v" + side].view(np.uint64))
            assert np.array_equal(t, g["t" + side])


def test_polygon_cuts_cover_the_polygon():
    """Closed forms of the ear clipping: a convex polygon becomes the fan
    from its first vertex, any simple polygon n - 2 triangles none of which
    turns against its winding, whose areas sum to its area."""
    v, t = dropin.load_obj(os.path.join(OBJ, "poly_pentagon.obj"))
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]
    for name, axes in [("poly_comb", (0, 1)), ("poly_ell", (0, 1)), ("poly_star_yz", (1, 2))]:
        for suffix in ("", "_cw"):
            v, t = dropin.load_obj(os.path.join(OBJ, name + suffix + ".obj"))
            p = v[:, axes]
            n = len(p)
            ring = np.arange(n) if suffix == "" else np.arange(n)[::-1]
            area = 0.5 * np.sum(p[ring, 0] * np.roll(p[ring], -1, 0)[:, 1] - np.roll(p[ring], -1, 0)[:, 0] * p[ring, 1])
            a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
            tri = 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
            assert len(t) == n - 2
            assert np.all(tri * area >= 0)  # ears of zero area are cut too
            assert abs(tri.sum() - area) < 1e-9 * abs(area)


def test_out_of_range_index_fails(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(IOError, match="outside"):
        dropin.load_obj(str(p))
    with pytest.raises(IOError, match="cannot open"):
        dropin.load_obj(str(tmp_path / "missing.obj"))


def test_export_reads_back(tmp_path):
    v, t = dropin.load_obj(os.path.join(OBJ, "cases", "complex", "b.obj"))
    out = tmp_path / "out.obj"
    dropin.export_obj(str(out), v, t)
    text = out.read_text().splitlines()
    assert text[0].startswith("v ") and len(text) == len(v) + len(t)
    assert text[len(v)] == "f " + " ".join(str(i + 1) for i in t[0])
    v2, t2 = dropin.load_obj(str(out))
    assert np.array_equal(t2, t)
    assert np.max(np.abs(v2 - v)) <= 5e-7 + 1e-6 * np.max(np.abs(v))
