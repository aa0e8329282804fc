"""Writes tests/golden/order_results.npz: the reference's combine() results in
the reference's own order, from the compiled reference (oracle/_ref/libsbref.so,
built from /root/reference by oracle/Makefile).

Run here, where the reference sources exist:
    make -C oracle ref && python tests/golden/gen_order_golden.py

Per case, for the cases where the reference's combine() succeeds:
* the result vertices in their order (resultVertices(), solidboolean.h), as
  SHA-256 of their little-endian FP64 bytes, and the array itself for small
  cases;
* the union, difference and intersection triangle lists in fetch order
  (fetchUnion / fetchDiff / fetchIntersect, src/solidboolean.cpp:512-565), as
  SHA-256 of their little-endian u64 indices, with their lengths;
* for the four shipped OBJ cases, the SHA-256 of the OBJ text the reference's
  test driver writes for them (test/main.cpp:73-158: union as is, difference
  shifted by (-2, 0, 0), intersection by (2, 0, 0), written with "v %f %f %f"
  and "f a b c", test/main.cpp:9-28).

Cases: the three subdivision-3 sphere pairs, the shipped OBJ cases, the 8x8
and 64x64-per-face cube-grid pairs shifted by (1e-6, 0.5, 0.5) (near-coplanar
faces, exact ties) and the subdivision-6 pairs of seeds 1 and 3.  The GPU box
never runs this script.
"""
from __future__ import annotations

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

SMALL_ARRAYS = 20000  # vertex / triangle counts up to which arrays are stored whole


def sha_f64(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def sha_u64(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, "<u8").tobytes()).hexdigest()


def merged_obj(v: np.ndarray, lists) -> bytes:
    """The driver's merged result: each list appended with its own copy of the
    vertices moved by its offset (vertex + offset in FP64), then every vertex
    and every face written as test/main.cpp:9-28 writes them."""
    out_v, out_f = [], []
    base = 0
    for tris, off in zip(lists, ((0.0, 0.0, 0.0), (-2.0, 0.0, 0.0), (2.0, 0.0, 0.0))):
        out_f.append(np.asarray(tris, np.int64) + base)
        out_v.append(v + np.array(off))
        base += len(v)
    lines = ["v %f %f %f\n" % (x, y, z) for x, y, z in np.concatenate(out_v).tolist()]
    lines += ["f %d %d %d\n" % (a + 1, b + 1, c + 1) for a, b, c in np.concatenate(out_f).tolist()]
    return "".join(lines).encode()


def record(L: RefLib, name: str, a: mg.Mesh, b: mg.Mesh, rec: dict, driver: bool):
    t0 = time.time()
    ra = L.mesh(a.vertices, a.triangles)
    rb = L.mesh(b.vertices, b.triangles)
    bo = RefBoolean(ra, rb)
    rec[f"{name}/ok"] = np.array(bool(bo.ok))
    if bo.ok:
        v = bo.vertices()
        lists = [bo.fetch(i).astype(np.int64) for i in range(3)]
        rec[f"{name}/vertices_sha"] = np.array(sha_f64(v))
        rec[f"{name}/n_vertices"] = np.array(len(v))
        if len(v) <= SMALL_ARRAYS:
            rec[f"{name}/vertices"] = v
        for k, what in enumerate(("union", "diff", "intersect")):
            rec[f"{name}/{what}_sha"] = np.array(sha_u64(lists[k]))
            rec[f"{name}/{what}_count"] = np.array(len(lists[k]))
            if len(lists[k]) <= SMALL_ARRAYS:
                rec[f"{name}/{what}"] = lists[k].astype(np.uint32)
        if driver:
            rec[f"{name}/merged_obj_sha"] = np.array(hashlib.sha256(merged_obj(v, lists)).hexdigest())
    print(f"  {name}: combine {'ok' if bo.ok else 'fails'}, {time.time() - t0:.1f}s", flush=True)


def main():
    L = RefLib()
    rec: dict = {}
    for seed in (1, 2, 3):
        a, b = mg.small_pair(seed)
        record(L, f"small_s{seed}", a, b, rec, False)
    for name in ("simple-ring", "cube-sphere", "complex", "addax-and-meerkat"):
        with np.load(os.path.join(HERE, f"case_{name}.npz")) as z:
            a, b = mg.Mesh(z["va"], z["ta"]), mg.Mesh(z["vb"], z["tb"])
        record(L, f"case_{name}", a, b, rec, True)
    for n in (8, 64):
        a, b = mg.cube_pair(n)
        record(L, f"cube_n{n}", a, b, rec, False)
    for seed in (1, 3):
        a, b = mg.icosphere_pair(6, seed, (0.5, 0.2, 0.1))
        record(L, f"sub6_s{seed}", a, b, rec, False)
    np.savez_compressed(os.path.join(HERE, "order_results.npz"), **rec)


if __name__ == "__main__":
    main()
