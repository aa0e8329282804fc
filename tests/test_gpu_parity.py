"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and the oracle.  Integer, index and FP64 results are compared bit for
bit; the FP32-filtered mode must give exactly the FP64 results."""
import numpy as np
import pytest

from conftest import FULL_CASES, golden_exists, load_golden
from oracle.oracle import OracleMesh
from solidboolean_amd import meshgen as mg
from solidboolean_amd import native as N

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


RAY_PRECISIONS = [N.FP64, N.FP32_FILTERED]


class ray_precision:
    """Runs the enclosed classification calls with the context's rays in the
    given precision (sbg_tuning.ray_precision), then restores the tuning."""

    def __init__(self, gpu, precision):
        self.gpu, self.precision = gpu, precision

    def __enter__(self):
        self.saved = self.gpu.tuning()
        self.gpu.set_tuning(**{**self.saved, "ray_precision": self.precision})

    def __exit__(self, *exc):
        self.gpu.set_tuning(**self.saved)


def sorted_pairs(p):
    p = np.asarray(p, np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def f32_down(x):
    """Largest float32 at or below each float64 value."""
    y = np.asarray(x, np.float64).astype(np.float32)
    return np.where(y.astype(np.float64) > x, np.nextafter(y, np.float32(-np.inf)), y)


def f32_up(x):
    """Smallest float32 at or above each float64 value."""
    y = np.asarray(x, np.float64).astype(np.float32)
    return np.where(y.astype(np.float64) < x, np.nextafter(y, np.float32(np.inf)), y)


def check_bvh_tight(nodes, root, rbox, boxes64):
    """Every stored child box is exactly the union of the FP32 outward leaf
    boxes below it (and the root box the union of all): the refit is exact,
    not merely covering, whatever tile or top node computed a box."""
    lo32, hi32 = f32_down(boxes64[:, :3]), f32_up(boxes64[:, 3:])
    if root & 0x80000000:
        t = root & 0x7FFFFFFF
        assert np.array_equal(rbox, np.concatenate([lo32[t], hi32[t]]))
        return
    order, stack = [], [root]
    while stack:
        r = stack.pop()
        order.append(r)
        for c in (int(nodes[r]["c0"]), int(nodes[r]["c1"])):
            if not c & 0x80000000:
                stack.append(c)
    union = {}
    for r in reversed(order):  # children before parents
        nd = nodes[r]
        u = None
        for c, stored in ((int(nd["c0"]), nd["box0"]), (int(nd["c1"]), nd["box1"])):
            if c & 0x80000000:
                t = c & 0x7FFFFFFF
                want = np.concatenate([lo32[t], hi32[t]])
            else:
                want = union[c]
            assert np.array_equal(stored, want), f"node {r}: child {c:#x} box {stored} != {want}"
            u = want if u is None else np.concatenate([np.minimum(u[:3], want[:3]), np.maximum(u[3:], want[3:])])
        union[r] = u
    assert np.array_equal(rbox, union[root])


def check_bvh(mesh, boxes64):
    """Every LBVH node's child boxes contain the boxes of the triangles below
    them (exactly their union), and every triangle is a leaf exactly once."""
    nodes, root, rbox = mesh.bvh()
    nt = mesh.nt
    if nt == 0:
        return
    lo = boxes64[:, :3]
    hi = boxes64[:, 3:]
    seen = np.zeros(nt, np.int64)
    stack = [(root, rbox)]
    while stack:
        ref, box = stack.pop()
        if ref & 0x80000000:
            t = ref & 0x7FFFFFFF
            seen[t] += 1
            assert np.all(box[:3] <= lo[t]) and np.all(box[3:] >= hi[t]), f"leaf {t} box not covered"
            continue
        nd = nodes[ref]
        for c, b in ((nd["c0"], nd["box0"]), (nd["c1"], nd["box1"])):
            assert np.all(b[:3] >= box[:3]) and np.all(b[3:] <= box[3:]), "child box outside parent box"
            stack.append((int(c), b))
    assert np.all(seen == 1)
    check_bvh_tight(nodes, root, rbox, boxes64)


@pytest.fixture(scope="module", params=FULL_CASES)
def case(request, gpu):
    g = load_golden(request.param)
    A = gpu.mesh(g["va"], g["ta"])
    B = gpu.mesh(g["vb"], g["tb"])
    return request.param, g, A, B


def test_prepare_normals_boxes_bvh(case):
    name, g, A, B = case
    assert np.array_equal(bits(A.normals()), bits(g["normals_a"])), name
    assert np.array_equal(bits(B.normals()), bits(g["normals_b"])), name
    assert np.array_equal(bits(A.boxes()), bits(g["boxes_a"][:, :6])), name
    assert np.array_equal(bits(B.boxes()), bits(g["boxes_b"][:, :6])), name
    check_bvh(A, g["boxes_a"][:, :6])
    check_bvh(B, g["boxes_b"][:, :6])


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_bvh_exact_across_refit_tiles(gpu, level):
    """Icospheres of 1,280 to 81,920 triangles: leaf counts that are not
    multiples of the refit's 1,024-leaf tiles, so ranges inside a tile, top
    nodes and a partial last tile all occur; every stored box is checked
    against the exact union of its leaves."""
    m = mg.radial_noise(mg.icosphere(level), 1e-3, mg.SplitMix64(level))
    mesh = gpu.mesh(m.vertices, m.triangles)
    corners = np.asarray(m.vertices, np.float64)[np.asarray(m.triangles, np.int64)]  # (nt, 3, 3)
    boxes64 = np.concatenate([corners.min(axis=1), corners.max(axis=1)], axis=1)
    assert np.array_equal(bits(mesh.boxes()), bits(boxes64))
    check_bvh(mesh, boxes64)


def test_pair_set(case):
    name, g, A, B = case
    p = A.find_pairs(B).array().astype(np.int64)
    assert np.array_equal(p, sorted_pairs(g["pairs"])), name  # set equal, canonical (a, b) order


@pytest.mark.parametrize("precision", [N.FP64, N.FP32_FILTERED])
def test_tri_tri(case, precision):
    name, g, A, B = case
    P = A.find_pairs(B)
    st = P.tri_tri(A, B, precision)
    status, seg = P.results()
    ref = sorted_pairs(g["pairs"])
    order = np.lexsort((g["pairs"][:, 1], g["pairs"][:, 0]))
    assert np.array_equal(status, g["status"][order]), name
    hit = status == 1
    ref_seg = np.zeros((len(g["pairs"]), 6))
    ref_seg[g["status"] == 1] = g["hit_seg"]
    assert np.array_equal(bits(seg[hit]), bits(ref_seg[order][hit])), name
    hp, hs = P.hits()
    assert np.array_equal(hp.astype(np.int64), ref[hit]) and np.array_equal(bits(hs), bits(ref_seg[order][hit]))
    assert st["tested"] == len(ref) and st["hits"] == int(hit.sum()) and st["coplanar"] == int((status == 2).sum())
    assert sum(st[k] for k in ("exit_plane1", "exit_plane2", "exit_interval", "hits", "coplanar")) == len(ref)
    if precision == N.FP32_FILTERED:
        assert st["filtered_pairs"] + st["hits"] + st["coplanar"] >= len(ref) - st["fallback_pairs"]


@pytest.mark.parametrize("rays", RAY_PRECISIONS)
def test_rays(case, gpu, rays):
    name, g, A, B = case
    with ray_precision(gpu, rays):
        axis_a, vote_a = A.classify_points(g["ray_pts_a"])
        axis_b, vote_b = B.classify_points(g["ray_pts_b"])
    assert np.array_equal(axis_a, g["ray_bits_a"]), name
    assert np.array_equal(axis_b, g["ray_bits_b"]), name
    assert np.array_equal(vote_a, (g["ray_bits_a"].sum(1) >= 2).astype(np.uint8))


@pytest.mark.parametrize("rays", RAY_PRECISIONS)
def test_classify_triangles(case, gpu, rays):
    name, g, A, B = case
    P = A.find_pairs(B)
    P.tri_tri(A, B, N.FP32_FILTERED, stats=False)
    with ray_precision(gpu, rays):
        assert np.array_equal(A.classify_triangles(B, P, 0), g["cls_a"]), name
        assert np.array_equal(B.classify_triangles(A, P, 1), g["cls_b"]), name
        ca, cb = A.classify_pair(B, P)
    assert np.array_equal(ca, g["cls_a"]), name
    assert np.array_equal(cb, g["cls_b"]), name
    if "side_a" in g and name != "cube_n8":
        # where the reference's combine() succeeds on a well-conditioned pair,
        # the per-triangle vote agrees with the side of the group each kept
        # triangle ended up in (on the near-coplanar cube pair the reference's
        # one-representative-per-group sides differ from per-triangle votes)
        keep = g["side_a"] != 2
        assert np.array_equal(g["cls_a"][keep], g["side_a"][keep]), name
        keep = g["side_b"] != 2
        assert np.array_equal(g["cls_b"][keep], g["side_b"][keep]), name


def test_gd_vectors_batch(gpu):
    g = load_golden("gd_vectors")
    status, seg = gpu.tri_tri_batch(g["a9"], g["b9"])
    assert np.array_equal(status, g["status"])
    assert np.array_equal(bits(seg[status == 1]), bits(g["seg"][g["status"] == 1]))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_oracle_random_seeds(gpu, seed):
    """Seeded subdivision-4 pairs with noise (oracle runs in about a second)."""
    a, b = mg.icosphere_pair(4, seed, (0.3, 0.2, -0.1), noise=1e-2, radius_b=0.9)
    OA, OB = OracleMesh(a.vertices, a.triangles), OracleMesh(b.vertices, b.triangles)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    ref = sorted_pairs(OA.find_pairs(OB))
    P = A.find_pairs(B)
    assert np.array_equal(P.array().astype(np.int64), ref)
    ost, oseg = OA.tri_tri_pairs(OB, ref)
    for prec in (N.FP64, N.FP32_FILTERED):
        P.tri_tri(A, B, prec)
        st, seg = P.results()
        assert np.array_equal(st, ost)
        assert np.array_equal(bits(seg[st == 1]), bits(oseg[ost == 1]))
    ia = np.zeros(a.nt, np.uint8)
    ia[ref[ost == 1, 0]] = 1
    ref_cls = OA.classify_triangles(OB, ia)
    for rays in RAY_PRECISIONS:
        with ray_precision(gpu, rays):
            assert np.array_equal(A.classify_triangles(B, P, 0), ref_cls), rays


def _rays_vs_oracle(A, OA, pts):
    """Per-axis ray answers and votes of points against a mesh, GPU vs oracle."""
    axis, vote = A.classify_points(pts)
    ref = np.stack([OA.points_inside(pts, k) for k in range(3)], 1)
    assert np.array_equal(axis, ref)
    assert np.array_equal(vote, (ref.sum(1) >= 2).astype(np.uint8))


def test_large_coordinates(gpu):
    """Meshes far from the origin: PositionKey components beyond 32 bits
    (|x| * 1e5 > 2^31), which the fast ray path hands to the wave path."""
    a, b = mg.icosphere_pair(3, 5, (0.3, 0.2, -0.1), noise=1e-2, radius_b=0.9)
    off = np.array([5.0e4, -3.0e4, 1.0e5])
    a = mg.Mesh(a.vertices * 1000.0 + off, a.triangles)
    b = mg.Mesh(b.vertices * 1000.0 + off, b.triangles)
    OA, OB = OracleMesh(a.vertices, a.triangles), OracleMesh(b.vertices, b.triangles)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    ref = sorted_pairs(OA.find_pairs(OB))
    P = A.find_pairs(B)
    assert np.array_equal(P.array().astype(np.int64), ref)
    P.tri_tri(A, B, N.FP32_FILTERED)
    st, _ = P.results()
    ost, _ = OA.tri_tri_pairs(OB, ref)
    assert np.array_equal(st, ost)
    ia = np.zeros(a.nt, np.uint8)
    ia[ref[ost == 1, 0]] = 1
    ib = np.zeros(b.nt, np.uint8)
    ib[ref[ost == 1, 1]] = 1
    for rays in RAY_PRECISIONS:
        with ray_precision(gpu, rays):
            assert np.array_equal(A.classify_triangles(B, P, 0), OA.classify_triangles(OB, ia))
            assert np.array_equal(B.classify_triangles(A, P, 1), OB.classify_triangles(OA, ib))
            pts = a.vertices[a.triangles].mean(1)[::7]
            _rays_vs_oracle(B, OB, pts)


def test_nested_shells_many_hits(gpu):
    """A target of four concentric shells (one mesh): rays cross up to eight
    faces, more distinct hits than the fast path keeps, so those rays finish
    in the wave path; votes and per-axis answers must still equal the oracle's."""
    shells = [mg.icosphere(3, r) for r in (0.4, 0.6, 0.8, 1.0)]
    V, T, base = [], [], 0
    for m in shells:
        V.append(m.vertices)
        T.append(m.triangles + base)
        base += m.vertices.shape[0]
    target = mg.Mesh(np.concatenate(V), np.concatenate(T))
    query = mg.transform(mg.icosphere(3, 0.7), None, (0.15, -0.05, 0.1))
    OT, OQ = OracleMesh(target.vertices, target.triangles), OracleMesh(query.vertices, query.triangles)
    Tm, Q = gpu.mesh(target.vertices, target.triangles), gpu.mesh(query.vertices, query.triangles)
    ref = sorted_pairs(OQ.find_pairs(OT))
    P = Q.find_pairs(Tm)
    assert np.array_equal(P.array().astype(np.int64), ref)
    P.tri_tri(Q, Tm, N.FP64)
    st, _ = P.results()
    iq = np.zeros(query.nt, np.uint8)
    iq[ref[st == 1, 0]] = 1
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.1, 1.1, (3000, 3))
    for rays in RAY_PRECISIONS:
        with ray_precision(gpu, rays):
            assert np.array_equal(Q.classify_triangles(Tm, P, 0), OQ.classify_triangles(OT, iq))
            _rays_vs_oracle(Tm, OT, pts)


def test_votes_into_page_locked_arrays(gpu):
    """Votes written into page-locked arrays (sbg_host_alloc) arrive by DMA
    and equal those staged into ordinary arrays, for classify_pair (both
    sides, two calls in a row with the arrays reused) and classify_triangles."""
    a, b = mg.icosphere_pair(5, 2, (0.3, 0.2, -0.1), noise=1e-2, radius_b=0.9)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    P = A.find_pairs(B)
    P.tri_tri(A, B, N.FP32_FILTERED, stats=False)
    ref_a, ref_b = A.classify_pair(B, P)
    out = (N.host_array(A.nt), N.host_array(B.nt))
    for _ in range(2):
        out[0][:] = 7
        out[1][:] = 7
        ca, cb = A.classify_pair(B, P, out=out)
        assert np.array_equal(ca, ref_a) and np.array_equal(cb, ref_b)
    t = N.host_array(B.nt)
    assert np.array_equal(B.classify_triangles(A, P, 1, out=t), ref_b)


def test_crowded_cells_stacked_copies(gpu):
    """Sixty stacked copies of one closed surface (a level-1 icosphere, 4,800
    triangles over the same 80 places): every cell is crowded, so the rays go
    to the wave kernel, each with sixty hits per face crossed, all at the same
    PositionKeys.  Per-axis answers and votes equal the oracle's, which counts
    each crossing once (distinct keys), at two per-lane cell limits and in
    both ray precisions, for scattered points and for clusters whose rays
    share cells."""
    s = mg.icosphere(1, 0.8)
    k = 60
    V = np.concatenate([s.vertices] * k)
    T = np.concatenate([s.triangles + j * s.vertices.shape[0] for j in range(k)])
    M, OM = gpu.mesh(V, T), OracleMesh(V, T)
    rng = np.random.default_rng(12)
    # scattered points, and clusters whose rays share cells
    pts = np.concatenate([rng.uniform(-1.0, 1.0, (400, 3)),
                          rng.uniform(-0.02, 0.02, (6, 40, 3)).reshape(-1, 3) + rng.uniform(-0.6, 0.6, (6, 1, 3)).repeat(40, 0).reshape(-1, 3)])
    saved = gpu.tuning()
    try:
        for rays in RAY_PRECISIONS:
            for cap in (16, 64):
                gpu.set_tuning(**{**saved, "ray_precision": rays, "ray_cell_cap": cap})
                _rays_vs_oracle(M, OM, pts)
    finally:
        gpu.set_tuning(**saved)


@pytest.mark.parametrize("name", ["sub6_s1", "sub6_s3"])
def test_sub6_digests(gpu, name):
    g = load_golden(name)
    seed = int(name[-1])
    a, b = mg.icosphere_pair(6, seed, (0.5, 0.2, 0.1))
    _check_digests(gpu, g, a, b)


def _check_digests(gpu, g, a, b):
    import hashlib

    def sha(*arrays):
        h = hashlib.sha256()
        for x in arrays:
            h.update(np.ascontiguousarray(x).tobytes())
        return h.hexdigest()

    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    P = A.find_pairs(B)
    p = P.array()
    assert len(p) == int(g["n_pairs"])
    assert sha(p.astype("<u4")) == str(g["pair_digest"])
    for prec in (N.FP64, N.FP32_FILTERED):
        st = P.tri_tri(A, B, prec)
        hp, hs = P.hits()
        assert sha(hp.astype("<u4"), np.ascontiguousarray(hs, "<f8")) == str(g["hits_digest"])
        assert st["coplanar"] == int(g["n_coplanar"])
    for rays in RAY_PRECISIONS:
        with ray_precision(gpu, rays):
            ca = A.classify_triangles(B, P, 0)
            cb = B.classify_triangles(A, P, 1)
            assert np.array_equal(np.bincount(ca, minlength=3), g["cls_count_a"]), rays
            assert np.array_equal(np.bincount(cb, minlength=3), g["cls_count_b"]), rays
            assert sha(ca) == str(g["cls_digest_a"]) and sha(cb) == str(g["cls_digest_b"]), rays
            pa, pb = A.classify_pair(B, P)
            assert sha(pa) == str(g["cls_digest_a"]) and sha(pb) == str(g["cls_digest_b"]), rays
    return P, st  # st: statistics of the FP32-filtered run


@pytest.mark.skipif(not golden_exists("large_s1"), reason="large fixture not generated")
def test_large_pair_digests(gpu):
    """1.31M-triangle noisy icosphere pair: pair set, every segment and every
    triangle's classification against the reference's digests."""
    g = load_golden("large_s1")
    a, b = mg.large_pair(1)
    _check_digests(gpu, g, a, b)


@pytest.mark.parametrize("name", ["stress_torus", "stress_cube"])
def test_stress_digests(gpu, name):
    """Filter stress cases: a 4.19M-triangle torus against a noisy
    subdivision-7 sphere, and two 512x512-per-face cube grids offset by
    (1e-6, 0.5, 0.5) with near-coplanar faces and exact-zero orientations:
    pair set, statuses and segments in both precisions, and every triangle's
    classification against the reference's digests."""
    g = load_golden(name)
    a, b = mg.torus_sphere_pair(1) if name == "stress_torus" else mg.cube_pair(512)
    assert mg.mesh_digest(a) == str(g["digest_a"]) and mg.mesh_digest(b) == str(g["digest_b"])
    _, st = _check_digests(gpu, g, a, b)
    assert st["hits"] == int(g["n_hits"])
    assert st["fallback_pairs"] + st["filtered_pairs"] <= st["tested"]


# ---------------------------------------------------------------------------
# edge cases
# ---------------------------------------------------------------------------
def test_empty_and_single_triangle(gpu):
    v = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    t = np.array([[0, 1, 2]])
    E = gpu.mesh(np.zeros((0, 3)), np.zeros((0, 3), np.int64))
    S = gpu.mesh(v, t)
    S2 = gpu.mesh(v * 2 - np.array([0.5, 0.5, 0.0]) + np.array([0, 0, 0.0]), t)
    assert len(E.find_pairs(S)) == 0 and len(S.find_pairs(E)) == 0
    P = S.find_pairs(S2)
    assert P.array().tolist() == [[0, 0]]
    P.tri_tri(S, S2, N.FP64)
    st, _ = P.results()
    assert st.tolist() == [2]  # coplanar: intersectTwoFaces returns false
    _, inside = S.classify_points(np.array([[0.2, 0.2, 0.0]]))
    assert inside.tolist() == [0]
    _, inside = E.classify_points(np.array([[0.2, 0.2, 0.0]]))
    assert inside.tolist() == [0]


def test_stream_ordering(gpu):
    """Meshes alternate between the context's two streams for their prepare;
    whatever is asked of a mesh right after its prepare (normals, boxes, tree,
    pairs, votes), with no synchronisation in between, equals the oracle's,
    step after step."""
    a, b = mg.icosphere_pair(4, 21, (0.3, 0.1, -0.2), noise=1e-2, radius_b=0.85)
    A = gpu.mesh(a.vertices, a.triangles, prepare=False)
    B = gpu.mesh(b.vertices, b.triangles, prepare=False)
    OA, OB = OracleMesh(a.vertices, a.triangles), OracleMesh(b.vertices, b.triangles)
    ref = sorted_pairs(OA.find_pairs(OB))
    ost, _ = OA.tri_tri_pairs(OB, ref)
    ia = np.zeros(a.nt, np.uint8)
    ia[ref[ost == 1, 0]] = 1
    ib = np.zeros(b.nt, np.uint8)
    ib[ref[ost == 1, 1]] = 1
    cls_a, cls_b = OA.classify_triangles(OB, ia), OB.classify_triangles(OA, ib)
    from oracle.oracle import normal
    nb = np.array([normal(*b.vertices[t]) for t in b.triangles])
    saved = gpu.tuning()
    for step in range(6):
        # the ray grids built at prepare time on the third stream (early_grids
        # 2), queued on the second stream beside the pair search (1), or built
        # when classification asks for them (0)
        gpu.set_tuning(early_grids=2 - step % 3)
        if step in (2, 3):
            # a search whose grids are still being built when the meshes
            # prepare again
            A.find_pairs(B)
        A.prepare()
        B.prepare()
        assert np.array_equal(bits(B.normals()), bits(nb)), step
        check_bvh(B, B.boxes())
        P = A.find_pairs(B)
        assert np.array_equal(P.array().astype(np.int64), ref), step
        P.tri_tri(A, B, N.FP32_FILTERED)
        ca, cb = A.classify_pair(B, P)
        assert np.array_equal(ca, cls_a) and np.array_equal(cb, cls_b), step
        assert np.array_equal(A.classify_triangles(B, P, 0), cls_a), step
    gpu.set_tuning(**saved)
    # two pair lists alive at once, classified in either order
    P1 = A.find_pairs(B)
    P1.tri_tri(A, B, N.FP32_FILTERED)
    P2 = A.find_pairs(B)
    P2.tri_tri(A, B, N.FP32_FILTERED)
    for Pk in (P1, P2, P2):
        ca, cb = A.classify_pair(B, Pk)
        assert np.array_equal(ca, cls_a) and np.array_equal(cb, cls_b)
    # a mesh made later lands on the other stream; its answers follow too
    C = gpu.mesh(b.vertices, b.triangles)
    P = A.find_pairs(C)
    assert np.array_equal(P.array().astype(np.int64), ref)


def test_disjoint_meshes(gpu):
    a, b = mg.icosphere_pair(3, 1, (5.0, 0.0, 0.0))
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    P = A.find_pairs(B)
    assert len(P) == 0
    P.tri_tri(A, B, N.FP32_FILTERED)
    assert np.all(A.classify_triangles(B, P, 0) == 0)


def test_identical_meshes_grow_queues(gpu):
    """A mesh against itself: every triangle pairs with all its neighbours
    (more candidates than the initial queue holds); shared vertices and
    self pairs put many orientations at or near zero (a triangle against
    itself is exactly coplanar only when its FP64 triple products round to 0)."""
    m = mg.icosphere(5)
    A, B = gpu.mesh(m.vertices, m.triangles), gpu.mesh(m.vertices, m.triangles)
    OA = OracleMesh(m.vertices, m.triangles)
    ref = sorted_pairs(OA.find_pairs(OracleMesh(m.vertices, m.triangles)))
    P = A.find_pairs(B)
    assert np.array_equal(P.array().astype(np.int64), ref)
    st = P.tri_tri(A, B, N.FP32_FILTERED)
    status, _ = P.results()
    ost, _ = OA.tri_tri_pairs(OracleMesh(m.vertices, m.triangles), ref)
    assert np.array_equal(status, ost)
    assert st["coplanar"] == int((ost == 2).sum()) > 0
    assert st["fallback_pairs"] > 0  # exact zeros cannot be certified in FP32


def test_degenerate_triangles(gpu):
    """Zero-area triangles: zero normal (Double::isZero) and exact handling."""
    v = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [0.5, 0.5, -1], [0.5, 0.5, 1], [0.6, 0.4, 1]])
    t = np.array([[0, 1, 2], [0, 1, 3], [1, 1, 3]])
    t2 = np.array([[4, 5, 6], [4, 4, 5]])
    A, B = gpu.mesh(v, t), gpu.mesh(v, t2)
    OA, OB = OracleMesh(v, t), OracleMesh(v, t2)
    from oracle.oracle import prepare

    n, _ = prepare(v, t)
    assert np.array_equal(bits(A.normals()), bits(n))
    ref = sorted_pairs(OA.find_pairs(OB))
    P = A.find_pairs(B)
    assert np.array_equal(P.array().astype(np.int64), ref)
    P.tri_tri(A, B, N.FP32_FILTERED)
    st, seg = P.results()
    ost, oseg = OA.tri_tri_pairs(OB, ref)
    assert np.array_equal(st, ost) and np.array_equal(bits(seg[st == 1]), bits(oseg[ost == 1]))


def test_host_pair_list_validation(gpu):
    a, b = mg.small_pair(1)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    bad = gpu.pairs(np.array([[0, b.nt]], np.uint32))
    with pytest.raises(N.SbgError):
        bad.tri_tri(A, B, N.FP64)
    with pytest.raises(N.SbgError):
        gpu.mesh(a.vertices, np.array([[0, 1, a.nv]]))
    # the message names the first offending triangle, not the largest index
    t = np.array(a.triangles, np.uint32)
    t[7, 2] = a.nv + 5
    t[3, 0] = a.nv
    with pytest.raises(N.SbgError, match=f"triangle 3 references vertex {a.nv} of {a.nv}"):
        gpu.mesh(a.vertices, t)
    # 64-bit indices go to sbg_mesh_create64, which refuses negative and wide
    # ones in the library (a negative int64 reads as a huge uint64)
    for dtype, bad in ((np.int64, -1), (np.int64, 2 ** 32), (np.uint64, 2 ** 40), (np.int64, a.nv)):
        t = np.array(a.triangles, dtype)
        t[5, 1] = bad
        with pytest.raises(N.SbgError, match=f"triangle 5 references vertex {int(np.uint64(t[5, 1]))} of {a.nv}"):
            gpu.mesh(a.vertices, t)


def test_staged_uploads_both_index_widths(gpu):
    """sbg_mesh_create (uint32 indices) and sbg_mesh_create64 (the reference's
    size_t indices, narrowed while staged) copy through page-locked staging on
    up to four host threads, each share cycling through two slots: a mesh of
    1,500,001 random triangles over 800,003 vertices (four shares per array,
    ragged ends) gives the oracle's normals and boxes bit for bit from both,
    and so does a one-triangle mesh (one share, one slot)."""
    from oracle.oracle import prepare

    rng = np.random.default_rng(11)
    for nv, nt in ((800_003, 1_500_001), (3, 1)):
        v = rng.standard_normal((nv, 3))
        t = rng.integers(0, nv, (nt, 3), dtype=np.int64)
        n, box = prepare(v, t)
        for tt in (t, t.astype(np.uint32)):
            M = gpu.mesh(v, tt)
            assert np.array_equal(bits(M.normals()), bits(n)), (nt, tt.dtype)
            assert np.array_equal(bits(M.boxes()), bits(box[:, :6])), (nt, tt.dtype)
            M.close()


def test_determinism(gpu):
    a, b = mg.icosphere_pair(5, 4, (0.3, 0.1, 0.2), noise=1e-3)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    p1 = A.find_pairs(B)
    p2 = A.find_pairs(B)
    assert np.array_equal(p1.array(), p2.array())
    p1.tri_tri(A, B, N.FP64)
    p2.tri_tri(A, B, N.FP32_FILTERED)
    s1, g1 = p1.results()
    s2, g2 = p2.results()
    assert np.array_equal(s1, s2) and np.array_equal(bits(g1[s1 == 1]), bits(g2[s2 == 1]))


def test_tuning_changes_speed_not_results(gpu):
    """Every setting of the single-pair knobs (breadth-first levels, depth-first
    waves, breadth-first only, ray-grid density, the cell size past which a
    ray gets a whole wave, when the grids are built) gives the same pairs,
    segments and classifications."""
    a, b = mg.icosphere_pair(5, 4, (0.3, -0.2, 0.1), noise=1e-3)
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    saved = gpu.tuning()

    def run():
        P = A.find_pairs(B)
        P.tri_tri(A, B, N.FP32_FILTERED, stats=False)
        st, seg = P.results()
        ca, cb = A.classify_pair(B, P)
        return (P.array(), st, bits(seg[st == 1]), A.classify_triangles(B, P, 0), B.classify_triangles(A, P, 1),
                ca, cb)

    try:
        ref = run()
        for kw in ({"bfs_levels": 0}, {"bfs_levels": 64, "dfs_waves": 64}, {"depth_first": 0},
                   {"grid_density": 1}, {"grid_density": 64}, {"dfs_waves": 16384, "bfs_levels": 4},
                   {"ray_cell_cap": 16}, {"ray_cell_cap": 4096, "grid_density": 1},
                   {"ray_cell_cap": 16, "ray_precision": N.FP64}, {"early_grids": 0}, {"early_grids": 1}):
            gpu.set_tuning(**{**saved, **kw})
            got = run()
            assert all(np.array_equal(x, y) for x, y in zip(ref, got)), kw
    finally:
        gpu.set_tuning(**saved)


def test_components_vs_oracle(gpu):
    """sbg_components against the oracle's restatement of the reference's face
    grouping: whole meshes, blocked bands, random blocked sets, several
    disjoint shells in one mesh, and the repeated-half-edge failure of a
    reversed triangle (none once that triangle is blocked)."""
    from oracle.oracle import face_groups

    rng = np.random.default_rng(9)
    m = mg.icosphere(3)
    c = m.vertices[m.triangles].mean(1)
    shells = [mg.icosphere(2, r) for r in (0.5, 1.0)] + [mg.transform(mg.icosphere(1), None, (3, 0, 0))]
    V, T, base = [], [], 0
    for s in shells:
        V.append(s.vertices)
        T.append(s.triangles + base)
        base += s.vertices.shape[0]
    multi = mg.Mesh(np.concatenate(V), np.concatenate(T))
    flipped = mg.Mesh(m.vertices, m.triangles.copy())
    flipped.triangles[17] = flipped.triangles[17][::-1]
    b17 = np.zeros(m.nt, bool)
    b17[17] = True
    cases = [(m, None), (m, np.abs(c[:, 2]) < 0.2), (m, rng.random(m.nt) < 0.3), (multi, None),
             (multi, rng.random(multi.nt) < 0.1), (flipped, b17 | (rng.random(m.nt) < 0.05))]
    for mesh, blocked in cases:
        G = gpu.mesh(mesh.vertices, mesh.triangles, prepare=False)
        got = G.components(None if blocked is None else blocked.astype(np.uint8)).astype(np.int64)
        got[got == 0xFFFFFFFF] = -1
        assert np.array_equal(got, face_groups(mesh.triangles, blocked))
    G = gpu.mesh(flipped.vertices, flipped.triangles, prepare=False)
    with pytest.raises(N.SbgError, match="repeated half-edge"):
        G.components()
    # the same context keeps working after the failure
    G = gpu.mesh(m.vertices, m.triangles, prepare=False)
    assert np.all(G.components() == 0)
    # at a large size, the closed form: one group per closed shell
    big = mg.icosphere(7)
    G = gpu.mesh(big.vertices, big.triangles, prepare=False)
    assert np.all(G.components() == 0)


def test_half_edges_vs_dict(gpu):
    """sbg_half_edges against a dictionary half-edge map built in index order
    (the reference's unordered_map keyed (from << 32) | to, first insert
    kept): opposite triangles, loop-edge owners, and the first triangle whose
    insertion meets a repeated directed edge."""
    def expect(T, queries):
        owner, first = {}, 0xFFFFFFFF
        for t, tri in enumerate(T):
            for k in range(3):
                key = (int(tri[k]) << 32) | int(tri[(k + 1) % 3])
                if key in owner:
                    first = min(first, t)
                else:
                    owner[key] = t
        opp = np.array([[owner.get((int(tri[(k + 1) % 3]) << 32) | int(tri[k]), 0xFFFFFFFF) for k in range(3)]
                        for tri in T], np.uint32).reshape(-1, 3)
        qo = np.array([owner.get(int(q), 0xFFFFFFFF) for q in queries], np.uint32)
        return opp, qo, first

    rng = np.random.default_rng(4)
    m = mg.icosphere(3)
    shells = mg.Mesh(np.concatenate([m.vertices, m.vertices + 3]), np.concatenate([m.triangles, m.triangles + m.nv]))
    T = shells.triangles.astype(np.uint32)
    q = np.array([(int(a) << 32) | int(b) for a, b in rng.integers(0, shells.nv, (50, 2))]
                 + [(int(T[i, 0]) << 32) | int(T[i, 1]) for i in range(0, len(T), 37)], np.uint64)
    cases = [(T, q), (T[rng.permutation(len(T))][:1000], q), (np.concatenate([T, T[5:6]]), q),
             (np.concatenate([T[:10], T[3:4][:, ::-1], T[10:]]), q), (np.array([[1, 1, 2], [3, 4, 5]], np.uint32), q),
             (np.zeros((0, 3), np.uint32), q)]
    for tris, queries in cases:
        opp, qo, first = gpu.half_edges(tris, queries)
        eo, eq, ef = expect(tris, queries)
        assert np.array_equal(opp, eo) and np.array_equal(qo, eq) and first == ef


@pytest.mark.parametrize("levels", [(4, 4), (1, 5), (5, 1)])
def test_pair_sort_runs(gpu, levels):
    """The quick pair sort (radix passes over the first triangle only, then
    each first triangle's run by the second) and its fallback when a run is
    too long (a coarse first mesh: every triangle meets hundreds of the fine
    one's): every search after the first on a pair queues it, and the list
    always equals the oracle's in (a, b) order."""
    la, lb = levels
    a = mg.icosphere(la, 1.0)
    b = mg.transform(mg.icosphere(lb, 0.9), None, (0.3, 0.2, -0.1))
    OA, OB = OracleMesh(a.vertices, a.triangles), OracleMesh(b.vertices, b.triangles)
    ref = sorted_pairs(OA.find_pairs(OB))
    A, B = gpu.mesh(a.vertices, a.triangles), gpu.mesh(b.vertices, b.triangles)
    runs = np.unique(ref[:, 0], return_counts=True)[1]
    for _ in range(3):
        P = A.find_pairs(B)
        assert np.array_equal(P.array().astype(np.int64), ref), (levels, int(runs.max()))
