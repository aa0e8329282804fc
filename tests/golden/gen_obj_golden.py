"""Writes the OBJ fixtures: OBJ texts that exercise the reader (number forms,
face index forms, polygons cut by the ear clipping, line endings, groups) under
tests/golden/obj/, the reference's own test cases copied from its test/cases
directory, and obj_golden.npz with what the reference's reader (tinyobjloader
as test/main.cpp:31-69 calls it, compiled from the reference tree into
oracle/_ref by oracle/Makefile) returns for each synthetic text.

Run in the development container (needs /root/reference and oracle/_ref):
    python tests/golden/gen_obj_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import reflib  # noqa: E402

OUT = os.path.join(HERE, "obj")
REF_CASES = "/root/reference/test/cases"
CASES = ["cube-sphere", "simple-ring", "complex", "addax-and-meerkat"]


def poly(name, pts, faces, extra=""):
    lines = [f"# {name}"] + [f"v {x!r} {y!r} {z!r}" for x, y, z in pts]
    lines += ["f " + " ".join(str(i + 1) for i in f) for f in faces]
    return "\n".join(lines) + "\n" + extra


def texts():
    rng = np.random.default_rng(5)
    t = {}
    # number forms, including more than 7 fraction digits and exponents
    nums = ["1e2", "-.5", "+3.25E-1", "0.123456789012", "-7.000000001", "12345678.9", "1.5e+3", "2.5E-3",
            "-0", ".5", "-.25e1", "1.", "3.14159265358979", "2.718281828459045", "1.4142135623730951",
            "abc", "7x", "1e", "-", "+.", "1e-45", "1e39", "0.30000000000000004", "123456789", "-1.0e-10"]
    rand = [f"{x:.{k}g}" for x, k in zip(rng.normal(0, 10, 300), rng.integers(1, 17, 300))]
    rand += [f"{x:.{k}f}" for x, k in zip(rng.uniform(-2, 2, 300), rng.integers(0, 16, 300))]
    allnum = nums + rand
    while len(allnum) % 3:
        allnum.append("0")
    lines = ["v " + " ".join(allnum[i:i + 3]) for i in range(0, len(allnum), 3)]
    lines += ["v 1", "v\t4\t5\t6", "  v 1 2 3 0.5 0.25 0.125", "v 7 8 9 # trailing"]
    t["numbers"] = "\n".join(lines) + "\n"
    # face index forms
    t["indices"] = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n"
                    "f 1/1/1 2/2/1 3/3/1\nf 1//1 2//1 4//1\nf 1/1 3/2 4/3\nf -4 -2 -1\nf -3/-2/-1 -2/-1/-1 -1/-3/-1\n"
                    "f 2 3 4\t\nf  1   3   2  \r\n")
    # polygons: convex and concave, either winding, in the three coordinate
    # planes and tilted, with collinear points, and degenerate ones
    ang = np.linspace(0, 2 * np.pi, 6)[:-1]
    penta = [(float(np.cos(a)), float(np.sin(a)), 0.0) for a in ang]
    hexa = [(float(np.cos(a)), 0.25, float(np.sin(a))) for a in np.linspace(0, 2 * np.pi, 7)[:-1]]
    ell = [(0, 0, 0), (2, 0, 0), (2, 1, 0), (1, 1, 0), (1, 2, 0), (0, 2, 0)]
    star = []
    for k in range(10):
        r = 1.0 if k % 2 == 0 else 0.4
        a = np.pi * k / 5
        star.append((0.1, float(r * np.cos(a)), float(r * np.sin(a))))
    dart = [(0, 0, 0), (2, 1, 0), (0, 2, 0), (0.5, 1, 0)]
    collinear = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (2, 1, 0), (1, 1, 0), (0, 1, 0)]
    flat = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)]
    rot = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    tilted = [tuple(float(c) for c in rot @ np.array(p, float)) for p in ell]
    comb = [(0, 0, 0), (5, 0, 0), (5, 3, 0), (4, 3, 0), (4, 1, 0), (3, 1, 0), (3, 3, 0), (2, 3, 0), (2, 1, 0),
            (1, 1, 0), (1, 3, 0), (0, 3, 0)]
    shapes = [("pentagon", penta), ("hexagon_xz", hexa), ("ell", ell), ("star_yz", star), ("dart", dart),
              ("collinear", collinear), ("flat", flat), ("tilted_ell", tilted), ("comb", comb)]
    for name, pts in shapes:
        n = len(pts)
        t[f"poly_{name}"] = poly(name, pts, [list(range(n))])
        t[f"poly_{name}_cw"] = poly(name, pts, [list(range(n))[::-1]])
    # a random closed-ish mesh of quads and triangles
    pts = [tuple(float(c) for c in rng.normal(0, 1, 3)) for _ in range(40)]
    faces = [list(rng.choice(40, size=int(rng.integers(3, 7)), replace=False)) for _ in range(60)]
    t["random_polygons"] = poly("random", pts, faces)
    # line endings, comments, short faces, groups with forward references
    t["lines_crlf"] = "# crlf\r\nv 0 0 0\r\nv 1 0 0\r\n\r\nv 1 1 0\r\nv 0 1 0\r\nf 1 2 3 4\r\nf 1 2\r\n"
    t["lines_cr"] = "v 0 0 0\rv 1 0 0\rv 1 1 0\rv 0 1 0\r  # comment\rf 4 3 2 1\r"
    t["groups_forward"] = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 3 4\ng second\nv 0 1 0\nv 0.5 1.5 0\n"
                           "f 1 2 3 5 4\no third\nf 5 4 1\ns 1\nusemtl none\nf 2 3 5\n")
    # failures of the reference's reader
    t["fail_zero_index"] = "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n"
    t["fail_word_index"] = "v 0 0 0\nv 1 0 0\nv 0 1 0\nf a b c\n"
    t["fail_zero_normal"] = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//0 2//1 3//1\n"
    return t


def main():
    if not reflib.available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    L = reflib.RefLib()
    os.makedirs(OUT, exist_ok=True)
    rec = {}
    for name, text in texts().items():
        path = os.path.join(OUT, name + ".obj")
        with open(path, "w", newline="") as f:
            f.write(text)
        try:
            v, tri = L.load_obj(path)
            rec[f"{name}__ok"] = np.array(1)
            rec[f"{name}__v"] = v
            rec[f"{name}__t"] = tri
        except IOError:
            rec[f"{name}__ok"] = np.array(0)
    np.savez_compressed(os.path.join(HERE, "obj_golden.npz"), **rec)
    for case in CASES:
        d = os.path.join(OUT, "cases", case)
        os.makedirs(d, exist_ok=True)
        for side in "ab":
            shutil.copyfile(os.path.join(REF_CASES, case, side + ".obj"), os.path.join(d, side + ".obj"))
    print(f"{len(rec)} arrays, {len(CASES)} reference cases")


if __name__ == "__main__":
    main()
