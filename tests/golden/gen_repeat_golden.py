"""Writes repeat_s1.npz: the subdivision-3 seed-1 pair with a copy of one of
mesh A's triangles appended to A (a directed edge that repeats, where the
reference's addUnintersectedTriangles, src/solidboolean.cpp:250-285, stops
copying that mesh's triangles) and the compiled reference's combine() result
on it.  Run in the development container (needs oracle/_ref):
    python tests/golden/gen_repeat_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", ".."))
import reflib  # noqa: E402
from solidboolean_amd import meshgen as mg  # noqa: E402


def main():
    L = reflib.RefLib()
    a, b = mg.small_pair(1)
    rec = {}
    for dup in (0, a.nt // 2, a.nt - 1):
        ta = np.concatenate([a.triangles, a.triangles[dup:dup + 1]])
        ra, rb = L.mesh(a.vertices, ta), L.mesh(b.vertices, b.triangles)
        bo = reflib.RefBoolean(ra, rb)
        rec[f"dup{dup}_ok"] = np.array(bo.ok)
        if bo.ok:
            rec[f"dup{dup}_face_counts"] = np.array([len(bo.fetch(i)) for i in range(3)])
            rec[f"dup{dup}_n_vertices"] = np.array(len(bo.vertices()))
        print(dup, bo.ok, [len(bo.fetch(i)) for i in range(3)] if bo.ok else None)
    rec["dups"] = np.array([0, a.nt // 2, a.nt - 1])
    np.savez_compressed(os.path.join(HERE, "repeat_s1.npz"), **rec)


if __name__ == "__main__":
    main()
