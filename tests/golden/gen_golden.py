"""Writes the golden fixtures under tests/golden/ from the compiled reference
(oracle/_ref/libsbref.so, built from /root/reference by oracle/Makefile).

Run here, where the reference sources exist:
    make -C oracle ref && python tests/golden/gen_golden.py [--small]

Every fixture is data: seeded inputs (or the digest that pins them) and the
reference's outputs on them.  The GPU box never runs this script.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from reflib import RefBoolean, RefLib  # noqa: E402

from solidboolean_amd import meshgen as mg  # noqa: E402

REF_CASES = "/root/reference/test/cases"
SHIPPED = ["simple-ring", "cube-sphere", "complex", "addax-and-meerkat"]


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def pair_digest(pairs: np.ndarray) -> str:
    p = np.asarray(pairs, np.int64).reshape(-1, 2)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return sha(p[order].astype("<u4"))


def hits_digest(pairs: np.ndarray, seg: np.ndarray) -> str:
    p = np.asarray(pairs, np.int64).reshape(-1, 2)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return sha(p[order].astype("<u4"), np.ascontiguousarray(seg[order], "<f8"))


def centroids(v: np.ndarray, t: np.ndarray) -> np.ndarray:
    # (v0 + v1 + v2) / 3.0, left to right (solidboolean.cpp:497-499)
    return (v[t[:, 0]] + v[t[:, 1]] + v[t[:, 2]]) / 3.0


def edge_midpoints(v: np.ndarray, t: np.ndarray) -> np.ndarray:
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]], 0)
    e = np.unique(np.sort(e, 1), axis=0)
    return (v[e[:, 0]] + v[e[:, 1]]) * 0.5


def votes(target, pts: np.ndarray):
    bits = np.stack([target.points_inside(pts, k)[0] for k in range(3)], 1)
    return bits, (bits.sum(1) >= 2).astype(np.uint8)


def case_record(L: RefLib, a: mg.Mesh, b: mg.Mesh, store_mesh: bool, full: bool, label: str):
    t0 = time.time()
    ra = L.mesh(a.vertices, a.triangles)
    rb = L.mesh(b.vertices, b.triangles)
    pairs, _ = ra.find_pairs(rb)
    status, raw, seg, _ = ra.tri_tri_pairs(rb, pairs)
    hit = np.nonzero(status == 1)[0]
    rec = dict(
        label=np.array(label),
        digest_a=np.array(mg.mesh_digest(a)),
        digest_b=np.array(mg.mesh_digest(b)),
        n_pairs=np.array(len(pairs)),
        pair_digest=np.array(pair_digest(pairs)),
        hits_digest=np.array(hits_digest(pairs[hit], seg[hit])),
        hit_pairs=pairs[hit].astype(np.uint32),
        hit_seg=seg[hit],
        n_coplanar=np.array(int((status == 2).sum())),
    )
    if store_mesh:
        rec.update(va=a.vertices, ta=a.triangles, vb=b.vertices, tb=b.triangles)
    if full:
        rec.update(
            pairs=pairs.astype(np.uint32),
            status=status,
            raw=raw,
            normals_a=ra.normals(),
            normals_b=rb.normals(),
            boxes_a=ra.boxes(),
            boxes_b=rb.boxes(),
            tree_a=np.array(list(ra.tree_stats().values())),
            tree_b=np.array(list(rb.tree_stats().values())),
        )
        # rays from points that sit on the other surface's vertices / edges
        qa = np.concatenate([b.vertices, edge_midpoints(b.vertices, b.triangles)], 0)
        qb = np.concatenate([a.vertices, edge_midpoints(a.vertices, a.triangles)], 0)
        rec["ray_pts_a"] = qa
        rec["ray_bits_a"], _ = votes(ra, qa)
        rec["ray_pts_b"] = qb
        rec["ray_bits_b"], _ = votes(rb, qb)
    # per-triangle centroid vote of every non-intersected triangle
    ia = np.zeros(a.nt, bool)
    ib = np.zeros(b.nt, bool)
    ia[pairs[hit, 0]] = True
    ib[pairs[hit, 1]] = True
    cls = []
    for (m, other, inter) in ((a, rb, ia), (b, ra, ib)):
        c = np.full(m.nt, 2, np.uint8)
        keep = np.nonzero(~inter)[0]
        _, v = votes(other, centroids(m.vertices, m.triangles)[keep])
        c[keep] = v
        cls.append(c)
    rec["cls_digest_a"] = np.array(sha(cls[0]))
    rec["cls_digest_b"] = np.array(sha(cls[1]))
    rec["cls_count_a"] = np.bincount(cls[0], minlength=3)
    rec["cls_count_b"] = np.bincount(cls[1], minlength=3)
    if full:
        rec["cls_a"] = cls[0]
        rec["cls_b"] = cls[1]
    bo = RefBoolean(ra, rb)
    rec["combine_ok"] = np.array(bo.ok)
    if bo.ok:
        rec["face_counts"] = np.array([len(bo.fetch(i)) for i in range(3)])
        rec["n_result_vertices"] = np.array(len(bo.vertices()))
        rec["side_a"] = bo.triangle_sides(0)
        rec["side_b"] = bo.triangle_sides(1)
        sa, _, sda = bo.groups(0)
        sb, _, sdb = bo.groups(1)
        rec["group_sizes_a"], rec["group_sides_a"] = sa, sda
        rec["group_sizes_b"], rec["group_sides_b"] = sb, sdb
        if full:
            v = bo.vertices()
            rec["result_vertices"] = v[np.lexsort((v[:, 2], v[:, 1], v[:, 0]))]
            for i, name in enumerate(("union", "diff", "intersect")):
                f = bo.fetch(i)
                rec[f"{name}_positions"] = np.sort(v[f].reshape(len(f), 9), axis=0) if len(f) else np.zeros((0, 9))
    print(f"  {label}: {len(pairs)} pairs, {len(hit)} hits, combine {'ok' if bo.ok else 'FAILS'}, "
          f"{time.time() - t0:.1f}s", flush=True)
    return rec


def gd_vectors(L: RefLib) -> dict:
    """Known-answer triangle pairs: random, near-touching, coplanar, exact-zero
    orientation cases; the reference's tri_tri_intersection_test_3d on each."""
    rng = mg.SplitMix64(12345)
    A, B = [], []

    def add(a, b):
        A.append(np.asarray(a, np.float64).reshape(9))
        B.append(np.asarray(b, np.float64).reshape(9))

    # random triangles in the unit cube
    u = rng.uniform(2000 * 18).reshape(2000, 18)
    for r in u:
        add(r[:9], r[9:])
    # small triangles near each other (exercise CONSTRUCT_INTERSECTION branches)
    c = rng.uniform(3000 * 3).reshape(3000, 3)
    d = (rng.uniform(3000 * 18).reshape(3000, 18) - 0.5) * 0.02
    for k in range(3000):
        add(np.tile(c[k], 3) + d[k, :9], np.tile(c[k], 3) + d[k, 9:])
    # integer-coordinate cases: exact zeros, shared vertices/edges, coplanar
    g = (rng.next_u64(4000 * 18) % np.uint64(5)).astype(np.float64).reshape(4000, 18) - 2.0
    for r in g:
        add(r[:9], r[9:])
    # coplanar (z = 0) and parallel-plane pairs
    z = rng.uniform(500 * 12).reshape(500, 12)
    for k in range(500):
        a = np.array([z[k, 0], z[k, 1], 0, z[k, 2], z[k, 3], 0, z[k, 4], z[k, 5], 0])
        b = np.array([z[k, 6], z[k, 7], 0, z[k, 8], z[k, 9], 0, z[k, 10], z[k, 11], 0])
        add(a, b)
        add(a, b + np.array([0, 0, 1e-12] * 3))
    # a vertex of B exactly on A's plane, crossing configurations
    for k in range(500):
        s = z[k]
        a = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float64)
        b = np.array([s[0], s[1], 0.0, s[2], s[3], 1.0, s[4], s[5], -1.0])
        add(a, b)
    A = np.array(A)
    B = np.array(B)
    n = len(A)
    ret = np.zeros(n, np.uint8)
    cop = np.zeros(n, np.uint8)
    seg = np.zeros((n, 6))
    for i in range(n):
        r, cp, s = L.tri_tri(A[i], B[i])
        ret[i], cop[i], seg[i] = r, cp, s
    # intersectTwoFaces status: 1 hit (non-coplanar), 2 coplanar, 0 none
    status = np.where(cop == 1, 2, np.where(ret == 1, 1, 0)).astype(np.uint8)
    seg[status != 1] = 0.0
    print(f"  gd vectors: {n} pairs, {int((status == 1).sum())} hits, {int((status == 2).sum())} coplanar")
    return dict(a9=A, b9=B, raw=ret, coplanar=cop, status=status, seg=seg)


def earcut_vectors(L: RefLib, n: int = 1500) -> dict:
    """Polygons with holes, collinear and duplicate points (the shapes the
    retriangulator hands over) and the vendored earcut's index lists."""
    rng = mg.SplitMix64(2024)

    def star(k, r0, r1, cx=0.0, cy=0.0):
        ang = np.sort(rng.uniform(k) * 2 * np.pi)
        rad = r0 + (r1 - r0) * rng.uniform(k)
        return np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], 1)

    xy, sizes, counts, idx = [], [], [], []
    for t in range(n):
        kind = t % 4
        k = 3 + int(rng.next_u64(1)[0] % np.uint64(60))
        if kind == 0:
            rings = [star(k, 0.5, 1.0)]
        elif kind == 1:
            rings = [star(k + 60, 0.7, 1.0), star(5, 0.05, 0.15, -0.3, 0.0), star(4, 0.05, 0.12, 0.3, 0.1)]
        elif kind == 2:
            tri = np.array([[0, 0], [1, 0], [0.3, 0.9]])
            pts = []
            for i in range(3):
                a, b = tri[i], tri[(i + 1) % 3]
                pts.append(a)
                for s in np.sort(rng.uniform(int(rng.next_u64(1)[0] % np.uint64(4)))):
                    pts.append(a + s * (b - a))
            rings = [np.array(pts), star(5, 0.02, 0.06, 0.4, 0.3)]
        else:
            g = (rng.next_u64(2 * k) % np.uint64(6)).astype(np.float64).reshape(k, 2)
            rings = [g]
        out = L.earcut(rings)
        for r in rings:
            xy.append(np.asarray(r, np.float64))
            sizes.append(len(r))
        counts.append(len(rings))
        idx.append(out)
    return dict(xy=np.concatenate(xy), ring_sizes=np.array(sizes), rings_per_case=np.array(counts),
                index_counts=np.array([len(i) for i in idx]), indices=np.concatenate(idx).astype(np.int64))


_FORK = {}


def _fork_votes(args):
    which, lo, hi = args
    target, pts = _FORK[which]
    return votes(target, pts[lo:hi])[1]


def parallel_votes(target, pts: np.ndarray, procs: int = 8) -> np.ndarray:
    """votes() split over forked workers that share `target` (the reference's
    prepared mesh) copy-on-write."""
    import multiprocessing as mp

    key = id(target)
    _FORK[key] = (target, pts)
    chunks = [(key, lo, min(lo + 65536, len(pts))) for lo in range(0, len(pts), 65536)]
    with mp.get_context("fork").Pool(procs) as pool:
        parts = pool.map(_fork_votes, chunks)
    del _FORK[key]
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8)


def stress_record(L: RefLib, a: mg.Mesh, b: mg.Mesh, label: str):
    """Digests of the candidate set, the intersecting pairs with their
    segments, and the per-triangle votes of both meshes; no combine()."""
    t0 = time.time()
    ra = L.mesh(a.vertices, a.triangles)
    rb = L.mesh(b.vertices, b.triangles)
    pairs, _ = ra.find_pairs(rb)
    status, _, seg, _ = ra.tri_tri_pairs(rb, pairs)
    hit = np.nonzero(status == 1)[0]
    rec = dict(label=np.array(label), digest_a=np.array(mg.mesh_digest(a)), digest_b=np.array(mg.mesh_digest(b)),
               n_pairs=np.array(len(pairs)), pair_digest=np.array(pair_digest(pairs)),
               hits_digest=np.array(hits_digest(pairs[hit], seg[hit])), n_hits=np.array(len(hit)),
               n_coplanar=np.array(int((status == 2).sum())),
               status_digest=np.array(sha(status[np.lexsort((pairs[:, 1], pairs[:, 0]))].astype(np.uint8))))
    for name, m, other, side in (("a", a, rb, 0), ("b", b, ra, 1)):
        inter = np.zeros(m.nt, bool)
        inter[pairs[hit, side]] = True
        c = np.full(m.nt, 2, np.uint8)
        keep = np.nonzero(~inter)[0]
        c[keep] = parallel_votes(other, centroids(m.vertices, m.triangles)[keep])
        rec[f"cls_digest_{name}"] = np.array(sha(c))
        rec[f"cls_count_{name}"] = np.bincount(c, minlength=3)
    print(f"  {label}: {len(pairs)} pairs, {len(hit)} hits, {int(rec['n_coplanar'])} coplanar, "
          f"{time.time() - t0:.1f}s", flush=True)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--small", action="store_true", help="skip the million-triangle cases")
    ap.add_argument("--stress", action="store_true", help="only the torus / cube stress digests")
    args = ap.parse_args()
    if args.stress:
        L = RefLib()
        print("torus (4,194,304 triangles) vs subdivision-7 sphere r=0.8 at (1, 0, 0), noise 1e-4")
        a, b = mg.torus_sphere_pair(1)
        np.savez_compressed(os.path.join(HERE, "stress_torus.npz"), **stress_record(L, a, b, "stress_torus"))
        print("cube grids 512x512 per face (3,145,728 triangles) vs the same shifted by (1e-6, 0.5, 0.5)")
        a, b = mg.cube_pair(512)
        np.savez_compressed(os.path.join(HERE, "stress_cube.npz"), **stress_record(L, a, b, "stress_cube"))
        return
    L = RefLib()
    out = HERE
    print("GD vectors")
    np.savez_compressed(os.path.join(out, "gd_vectors.npz"), **gd_vectors(L))
    print("earcut vectors")
    np.savez_compressed(os.path.join(out, "earcut_vectors.npz"), **earcut_vectors(L))
    print("subdivision-3 icosphere pairs")
    for seed in (1, 2, 3):
        a, b = mg.small_pair(seed)
        np.savez_compressed(os.path.join(out, f"small_s{seed}.npz"),
                            **case_record(L, a, b, True, True, f"small_s{seed}"))
    print("shipped cases")
    for name in SHIPPED:
        va, ta = L.load_obj(os.path.join(REF_CASES, name, "a.obj"))
        vb, tb = L.load_obj(os.path.join(REF_CASES, name, "b.obj"))
        rec = case_record(L, mg.Mesh(va, ta), mg.Mesh(vb, tb), True, True, name)
        np.savez_compressed(os.path.join(out, f"case_{name}.npz"), **rec)
    print("8x8-grid cube pair shifted by (1e-6, 0.5, 0.5): exact ties, near-coplanar faces")
    a, b = mg.cube_pair(8)
    np.savez_compressed(os.path.join(out, "cube_n8.npz"), **case_record(L, a, b, True, True, "cube_n8"))
    print("subdivision-6 icosphere pairs")
    for seed in (1, 3):
        a, b = mg.icosphere_pair(6, seed, (0.5, 0.2, 0.1))
        np.savez_compressed(os.path.join(out, f"sub6_s{seed}.npz"),
                            **case_record(L, a, b, False, False, f"sub6_s{seed}"))
    if args.small:
        return
    print("subdivision-8 noisy icosphere pair (1.31M triangles each)")
    a, b = mg.large_pair(1)
    np.savez_compressed(os.path.join(out, "large_s1.npz"), **case_record(L, a, b, False, False, "large_s1"))


if __name__ == "__main__":
    main()
