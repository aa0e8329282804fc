"""The drop-in SolidMesh / SolidBoolean C++ API end to end on the GPU against
the reference's combine() on the same inputs: success / failure, face counts
of union / difference / intersection, the side of every kept triangle, and the
result vertices (within 1e-12 relative)."""
import os

import numpy as np
import pytest

from conftest import FULL_CASES, GOLDEN, load_golden
from solidboolean_amd import dropin
from solidboolean_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def run(va, ta, vb, tb, reference_pair_order=False):
    a = dropin.SolidMesh(va, ta)
    b = dropin.SolidMesh(vb, tb)
    a.prepare()
    b.prepare()
    sb = dropin.SolidBoolean(a, b, reference_pair_order)
    ok = sb.combine()
    return ok, sb, (a, b)


def sorted_rows(v):
    return v[np.lexsort((v[:, 2], v[:, 1], v[:, 0]))]


@pytest.mark.parametrize("name", FULL_CASES)
def test_combine_matches_reference(name):
    g = load_golden(name)
    ok, sb, keep = run(g["va"], g["ta"], g["vb"], g["tb"])
    assert ok == bool(g["combine_ok"]), name
    if not ok:
        return
    if name == "cube_n8":
        # Near-coplanar faces: the reference's group sides here depend on which
        # triangle comes first in each group (its per-triangle votes disagree
        # within a group), which follows the order intersecting pairs are
        # processed in.  In (a, b) order that triangle differs;
        # test_combine_bit_exact_in_reference_order covers this case in the
        # reference's pair order.
        pytest.xfail("group representative follows the pair order; see the reference-order test")
    counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
    assert counts == list(g["face_counts"]), name
    assert np.array_equal(sb.triangle_sides(0), g["side_a"]), name
    assert np.array_equal(sb.triangle_sides(1), g["side_b"]), name
    v = sb.result_vertices()
    assert len(v) == int(g["n_result_vertices"]), name
    ref = g["result_vertices"]
    got = sorted_rows(v)
    scale = np.maximum(np.abs(ref), 1.0)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale), name


@pytest.mark.parametrize("name", [n for n in FULL_CASES if n != "cube_n8"])
def test_combine_in_reference_pair_order(name):
    """Welding in the reference's search order: the same faces, sides and
    points (to 1e-12), and at least as many points bit-equal to the
    reference's as in (a, b) order.  Within a cut triangle the kept point of
    a PositionKey cell is then the reference's; across triangles the
    reference's kept point follows its hash-map iteration, which neither
    order reproduces, so a few points stay one ulp apart."""
    g = load_golden(name)
    ref = g["result_vertices"]
    exact = []
    for order in (False, True):
        ok, sb, keep = run(g["va"], g["ta"], g["vb"], g["tb"], reference_pair_order=order)
        assert ok == bool(g["combine_ok"]), name
        if not ok:
            return
        counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
        assert counts == list(g["face_counts"]), name
        assert np.array_equal(sb.triangle_sides(0), g["side_a"]), name
        got = sorted_rows(sb.result_vertices())
        assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1.0)), name
        exact.append(int(np.all(got == ref, 1).sum()))
    assert exact[1] >= exact[0], (name, exact)


ORDER_CASES = ["small_s1", "small_s2", "small_s3", "case_simple-ring", "case_cube-sphere", "case_complex",
               "case_addax-and-meerkat", "cube_n8", "cube_n64", "sub6_s1", "sub6_s3"]


def order_case(name):
    if name == "cube_n64":
        return mg.cube_pair(64)
    if name.startswith("sub6_s"):
        return mg.icosphere_pair(6, int(name[-1]), (0.5, 0.2, 0.1))
    g = load_golden(name)
    return mg.Mesh(g["va"], g["ta"]), mg.Mesh(g["vb"], g["tb"])


def sha_bytes(a, dtype):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ORDER_CASES)
def test_combine_bit_exact_in_reference_order(name):
    """In the reference's pair order the drop-in's results are the
    reference's bit for bit and in the same order: every result vertex
    (numbering included), and the union, difference and intersection lists
    triangle by triangle, written by the compiled reference
    (tests/golden/gen_order_golden.py).  cube_n8 / cube_n64 have
    near-coplanar faces and exact ties."""
    with np.load(os.path.join(GOLDEN, "order_results.npz")) as z:
        g = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}
    a, b = order_case(name)
    ok, sb, keep = run(a.vertices, a.triangles, b.vertices, b.triangles, reference_pair_order=True)
    assert ok == bool(g["ok"]), name
    if not ok:
        return
    v = sb.result_vertices()
    assert len(v) == int(g["n_vertices"]), name
    if "vertices" in g:
        assert np.array_equal(v.view(np.uint64), g["vertices"].view(np.uint64)), name
    assert sha_bytes(v, "<f8") == str(g["vertices_sha"]), name
    for what, got in (("union", sb.fetch_union()), ("diff", sb.fetch_diff()), ("intersect", sb.fetch_intersect())):
        assert len(got) == int(g[f"{what}_count"]), (name, what)
        if what in g:
            assert np.array_equal(got, g[what].astype(np.int64)), (name, what)
        assert sha_bytes(got, "<u8") == str(g[f"{what}_sha"]), (name, what)


@pytest.mark.parametrize("seed", [1, 3])
def test_combine_subdivision6(seed):
    g = load_golden(f"sub6_s{seed}")
    a, b = mg.icosphere_pair(6, seed, (0.5, 0.2, 0.1))
    ok, sb, keep = run(a.vertices, a.triangles, b.vertices, b.triangles)
    assert ok == bool(g["combine_ok"])
    if ok:
        counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
        assert counts == list(g["face_counts"])
        assert len(sb.result_vertices()) == int(g["n_result_vertices"])
        assert np.array_equal(sb.triangle_sides(0), g["side_a"])
        assert np.array_equal(sb.triangle_sides(1), g["side_b"])


def test_fetch_appends_and_diff_reverses():
    g = load_golden("small_s1")
    ok, sb, keep = run(g["va"], g["ta"], g["vb"], g["tb"])
    assert ok
    u = sb.fetch_union()
    i = sb.fetch_intersect()
    d = sb.fetch_diff()
    # union = A-out + B-out, intersection = A-in + B-in, difference = A-out +
    # reversed B-in (src/solidboolean.cpp:512-565)
    k = 0
    while k < min(len(u), len(d)) and np.array_equal(u[k], d[k]):
        k += 1
    b_in = len(d) - k
    assert b_in > 0 and k > 0
    assert np.array_equal(d[k:], i[len(i) - b_in:][:, ::-1])


def signed_volume(v, t):
    t = np.asarray(t, np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def watertight(t):
    """Every directed edge once, and its reverse present: a closed, consistently
    oriented surface."""
    t = np.asarray(t, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    fwd = e[:, 0] * (1 << 32) + e[:, 1]
    u, c = np.unique(fwd, return_counts=True)
    return bool(np.all(c == 1) and np.all(np.isin(e[:, 1] * (1 << 32) + e[:, 0], u)))


@pytest.mark.parametrize("name", [n for n in FULL_CASES if n != "cube_n8"])
def test_robust_weld_agrees_where_reference_succeeds(name):
    g = load_golden(name)
    a = dropin.SolidMesh(g["va"], g["ta"])
    b = dropin.SolidMesh(g["vb"], g["tb"])
    a.prepare()
    b.prepare()
    sb = dropin.SolidBoolean(a, b, robust_weld=True)
    assert sb.combine() == bool(g["combine_ok"])
    counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
    assert counts == list(g["face_counts"]), name
    assert np.array_equal(sb.triangle_sides(0), g["side_a"]), name
    assert np.array_equal(sb.triangle_sides(1), g["side_b"]), name


@pytest.mark.parametrize("which", ["sub6_s3", "large_s1"])
def test_robust_weld_succeeds_where_reference_fails(which):
    """Pairs on which the reference's combine() fails ("Attach point to
    triangle edge failed"), as ours does by default: with the robust weld the
    union, difference and intersection are closed, consistently oriented
    surfaces, and their volumes obey inclusion-exclusion:
    V(A u B) + V(A n B) = V(A) + V(B) and V(A - B) = V(A) - V(A n B)."""
    g = load_golden(which)
    assert not bool(g["combine_ok"])
    pa, pb = mg.icosphere_pair(6, 3, (0.5, 0.2, 0.1)) if which == "sub6_s3" else mg.large_pair(1)
    assert mg.mesh_digest(pa) == str(g["digest_a"]) and mg.mesh_digest(pb) == str(g["digest_b"])
    a = dropin.SolidMesh(pa.vertices, pa.triangles)
    b = dropin.SolidMesh(pb.vertices, pb.triangles)
    a.prepare()
    b.prepare()
    assert not dropin.SolidBoolean(a, b).combine()
    sb = dropin.SolidBoolean(a, b, robust_weld=True)
    assert sb.combine()
    v = sb.result_vertices()
    u, d, i = sb.fetch_union(), sb.fetch_diff(), sb.fetch_intersect()
    assert watertight(u) and watertight(d) and watertight(i)
    va, vb = signed_volume(pa.vertices, pa.triangles), signed_volume(pb.vertices, pb.triangles)
    vu, vd, vi = signed_volume(v, u), signed_volume(v, d), signed_volume(v, i)
    assert vi > 0 and vu > max(va, vb)
    assert abs(vu + vi - (va + vb)) < 1e-9 * (va + vb)
    assert abs(vd - (va - vi)) < 1e-9 * va


def test_repeated_half_edge_falls_back_to_host_maps():
    """A directed edge that repeats (a copy of one of A's triangles appended):
    the GPU half-edge tables find it, and the host maps with the reference's
    insertion-order rules give the compiled reference's face counts."""
    g = load_golden("repeat_s1")
    a, b = mg.small_pair(1)
    for dup in g["dups"]:
        ta = np.concatenate([a.triangles, a.triangles[dup:dup + 1]])
        ok, sb, keep = run(a.vertices, ta, b.vertices, b.triangles)
        assert ok == bool(g[f"dup{dup}_ok"]), dup
        if ok:
            counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
            assert counts == list(g[f"dup{dup}_face_counts"]), dup
            assert len(sb.result_vertices()) == int(g[f"dup{dup}_n_vertices"]), dup


@pytest.mark.parametrize("name", FULL_CASES)
def test_robust_weld_succeeds_where_default_does(name):
    """The robust weld's chain (crossings, then PositionKey cells with edge
    hints, then plain PositionKey cells) succeeds wherever the default weld
    does; on the near-coplanar cube pair (exact ties) the crossings cannot
    close the loops and a later stage gives the result."""
    g = load_golden(name)
    oks = []
    for robust in (False, True):
        a = dropin.SolidMesh(g["va"], g["ta"])
        b = dropin.SolidMesh(g["vb"], g["tb"])
        a.prepare()
        b.prepare()
        oks.append(dropin.SolidBoolean(a, b, robust_weld=robust).combine())
    assert oks[1] or not oks[0]


def test_robust_weld_torus_stress_pair():
    """The torus x sphere stress pair (4,194,304 + 327,680 triangles), where
    the reference's combine() and the default weld fail: the robust chain
    succeeds (its PositionKey-cell stage with edge hints), and the volumes of
    union, intersection and difference agree with inclusion-exclusion."""
    a, b = mg.torus_sphere_pair(1)
    A = dropin.SolidMesh(a.vertices, a.triangles)
    B = dropin.SolidMesh(b.vertices, b.triangles)
    A.prepare()
    B.prepare()
    assert not dropin.SolidBoolean(A, B).combine()
    sb = dropin.SolidBoolean(A, B, robust_weld=True)
    assert sb.combine()
    v = sb.result_vertices()
    va, vb = signed_volume(a.vertices, a.triangles), signed_volume(b.vertices, b.triangles)
    vu, vd, vi = (signed_volume(v, f) for f in (sb.fetch_union(), sb.fetch_diff(), sb.fetch_intersect()))
    assert vi > 0 and vu > max(va, vb)
    assert abs(vu + vi - (va + vb)) < 1e-7 * (va + vb)
    assert abs(vd - (va - vi)) < 1e-7 * va


def test_threads_share_the_gpu():
    """Distinct SolidBoolean instances on different host threads (the
    reference allows it): each combine() gives what it gives alone."""
    import threading

    names = ["small_s1", "small_s2", "small_s3", "case_complex"]
    expect = {}
    for n in names:
        g = load_golden(n)
        ok, sb, keep = run(g["va"], g["ta"], g["vb"], g["tb"])
        expect[n] = (ok, [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())])
    got, errors = {}, []

    def work(n, k):
        try:
            g = load_golden(n)
            for r in range(3):
                ok, sb, keep = run(g["va"], g["ta"], g["vb"], g["tb"])
                got[(n, k, r)] = (ok, [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())])
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(n, k)) for k in range(2) for n in names]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    assert len(got) == 2 * 3 * len(names)
    for (n, k, r), v in got.items():
        assert v == expect[n], (n, k, r)
