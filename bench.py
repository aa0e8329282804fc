"""Benchmark of the boolean hot path on one MI355X.

Workload: the 1.31M-triangle pair - two subdivision-8 icospheres (1,310,720
triangles each, radius 1) with seeded radial noise 1e-3, the second rotated by
a seeded random rotation and offset by (0.4, 0.25, 0.1).

One step = the whole GPU hot path on mesh data already in HBM:
  prepare both meshes (records, normals, leaf boxes, Morton sort, LBVH),
  tree-vs-tree candidate search, Guigue-Devillers on every candidate pair
  (FP32-filtered), and the inside/outside vote of every non-intersected
  triangle of both meshes (x and y rays, z where they disagree: the same
  2-of-3 vote as three rays), results copied to the host.

`value` is the metric's first number: candidate tri-pair tests per second of
the Guigue-Devillers kernel (FP32-filtered), i.e. candidates per step divided
by the kernel's HIP-event time in a repeat of the timed steps with stage
events on.  `ms_per_step` is the second: end-to-end hot-path milliseconds,
wall clock over the timed steps (no events recorded).  The plain-FP64 step
(FP64 Guigue-Devillers and FP64 ray tests) and the filtered batched step, the
saturated (2^24-test) kernel rate, the stage breakdown, the roofline of the
step's dominant kernel (`roofline`: the x/y ray kernel, timed alone with HIP
events) and of the GD kernel over the 4,096-pair batch's candidates
(`batch.roofline_gd`), the whole boolean through the drop-in, and the CPU
baseline (the compiled reference, or the oracle port, on this host) are
reported in the same JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6

# algorithmic bytes in the units of the measurement plan: a triangle's
# coordinates are 72 B in FP64 and 36 B in FP32 (the records the kernels read
# are 96 B and 48 B: a normal or error-bound pads ride along)
TRI64_B = 72
TRI32_B = 36
PAIR_B = 8
STATUS_B = 1
SEG_B = 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


STEP_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kernel_trace_and_traffic_1310720_pair_step.md")
RAYS_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_pmc_rays_1310720_pair.md")
GD_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_pmc_gd_4096_batch.md")
# the per-lane cell limit (sbg_tuning.ray_cell_cap) that summary was taken at:
# the library default, which `bench.py --rays-only` runs with
RAYS_PROFILE_CELL_CAP = 64

# algorithmic bytes of the x/y ray kernel (k_rays_grid) in the units of the
# ray-parity row of the measurement plan: per ray its 24-B query point and
# 1-B result; per candidate visited (a scanned grid entry, which stands in for
# a node visit) its 8 B of box codes; per triangle tested its coordinates,
# 72 B in FP64 or 36 B in FP32
RAY_QUERY_B = 24 + 1
RAY_ENTRY_B = 8
RAY_TEST64_B = 72
RAY_TEST32_B = 36


def ray_bytes(work: dict, filtered: bool) -> float:
    return (work["rays"] * RAY_QUERY_B + work["scanned"] * RAY_ENTRY_B
            + work["passed"] * (RAY_TEST32_B if filtered else RAY_TEST64_B))


def measured_traffic(kernels, grid, profile: str = STEP_TRAFFIC_PROFILE):
    """HBM bytes per launch of `kernels` from a committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by the copy kernel's
    calibration), or None when the summary does not cover them.  `grid` is
    the launch's work-item count, or None for each kernel's smallest profiled
    launch (a lone-kernel launch over the exact item count, where the
    profiled program also launches it over a larger capacity)."""
    try:
        rows = {}
        for line in open(profile):
            cells = [c.strip().strip("`") for c in line.split("|")]
            if len(cells) == 7 and cells[2].isdigit():
                rows[(cells[1].replace("sbg::", ""), int(cells[2]))] = float(cells[5]) * 1e6
        total = 0.0
        for k in kernels:
            g = grid if grid is not None else min(gr for (name, gr) in rows if name == k)
            total += rows[(k, g)]
        return total
    except (OSError, KeyError, ValueError):
        return None


def gd_bytes(n: int, hits: int, fallback: int, precision: int) -> float:
    """Algorithmic HBM bytes of one Guigue-Devillers launch over n pairs."""
    if precision == 0:
        return n * (2 * TRI64_B + PAIR_B + STATUS_B) + hits * SEG_B
    return n * (2 * TRI32_B + PAIR_B + STATUS_B) + (hits + fallback) * 2 * TRI64_B + hits * SEG_B


def cpu_baseline(a, b, pairs_ref, budget_s: float = 12.0):
    """The reference's own code (oracle/_ref/libsbref.so, built from its sources)
    when present, else the oracle's C port: single core, same workload, bounded."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    kind = "reference"
    try:
        import reflib

        if not reflib.available():
            raise ImportError
        L = reflib.RefLib()
        ra, rb = L.mesh(a.vertices, a.triangles), L.mesh(b.vertices, b.triangles)
        prep_s = ra.prepare_seconds + rb.prepare_seconds
        pairs, search_s = ra.find_pairs(rb)
        gd_s = []
        t_end = time.perf_counter() + min(2.0, budget_s / 4)
        while time.perf_counter() < t_end or not gd_s:
            status, _, _, s = ra.tri_tri_pairs(rb, pairs)
            gd_s.append(s)
        gd_s = float(np.median(gd_s))
        hit = status == 1

        def classify(mesh, other, inter, n_sample):
            keep = np.nonzero(~inter)[0]
            sample = keep if n_sample is None else keep[:: max(1, len(keep) // n_sample)]
            c = (mesh.vertices[mesh.triangles[sample, 0]] + mesh.vertices[mesh.triangles[sample, 1]]
                 + mesh.vertices[mesh.triangles[sample, 2]]) / 3.0
            t = 0.0
            for k in range(3):
                _, s = other.points_inside(c, k)
                t += s
            return t / len(sample) * len(keep), len(sample)
    except ImportError:
        kind = "port"
        from oracle.oracle import OracleMesh

        t0 = time.perf_counter()
        ra, rb = OracleMesh(a.vertices, a.triangles), OracleMesh(b.vertices, b.triangles)
        prep_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        pairs = ra.find_pairs(rb)
        search_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        status, _ = ra.tri_tri_pairs(rb, pairs)
        gd_s = time.perf_counter() - t0
        hit = status == 1

        def classify(mesh, other, inter, n_sample):
            keep = np.nonzero(~inter)[0]
            sample = keep if n_sample is None else keep[:: max(1, len(keep) // n_sample)]
            c = (mesh.vertices[mesh.triangles[sample, 0]] + mesh.vertices[mesh.triangles[sample, 1]]
                 + mesh.vertices[mesh.triangles[sample, 2]]) / 3.0
            t0 = time.perf_counter()
            other.classify_points(c)
            return (time.perf_counter() - t0) / len(sample) * len(keep), len(sample)

    ia = np.zeros(a.nt, bool)
    ib = np.zeros(b.nt, bool)
    ia[pairs[hit, 0]] = True
    ib[pairs[hit, 1]] = True
    # every non-intersected triangle of both meshes with the reference's code
    # (about 11 s on one core); the oracle port, much slower, is sampled
    n_sample = None if kind == "reference" else 20000
    cls_a, sa = classify(a, rb, ia, n_sample)
    cls_b, sb = classify(b, ra, ib, n_sample)
    total_s = prep_s + search_s + gd_s + cls_a + cls_b
    return {
        "value": len(pairs) / gd_s,
        "unit": "tri-pair tests/s",
        "cores": 1,
        "kind": kind,
        "end_to_end_ms": total_s * 1e3,
        "stage_ms": {"prepare": prep_s * 1e3, "search": search_s * 1e3, "tri_tri": gd_s * 1e3,
                     "classify": (cls_a + cls_b) * 1e3},
        "sample": (f"same 1.31M-triangle pair on this host, 1 thread: full prepare of both meshes, full candidate "
                   f"search, Guigue-Devillers over all {len(pairs)} candidates (median of repeats), and the "
                   f"3-ray vote of {sa}+{sb} non-intersected triangles"
                   + (" (all of them)" if n_sample is None else " sampled, scaled to all of them")),
        "host_cpu": _cpu_model(),
        "nproc": os.cpu_count(),
    }


def fp64_vs_batched(ctx, a, b, step, N, steps: int, graph: bool = True) -> dict:
    """The 1.31M-triangle pair end to end three ways: the plain FP64 path
    (single-pair API, FP64 Guigue-Devillers and FP64 ray tests), the
    FP32-filtered single-pair path, and the FP32-filtered batched path (the
    batch API with this one pair, replaying its captured graph)."""
    def timed(precision):
        saved = ctx.tuning()
        ctx.set_tuning(**{**saved, "ray_precision": precision})
        for _ in range(2):
            step(precision)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(precision)
        ctx.synchronize()
        ctx.set_tuning(**saved)
        return (time.perf_counter() - t0) / steps * 1e3

    plain = timed(N.FP64)
    single = timed(N.FP32_FILTERED)
    batch = ctx.batch([(a, b)])
    if not graph:
        batch.set_tuning(use_graph=0)
    batch.run(N.FP32_FILTERED)
    batch.run(N.FP32_FILTERED)
    dev, t0 = [], time.perf_counter()
    for _ in range(steps):
        dev.append(batch.run(N.FP32_FILTERED))
    wall = (time.perf_counter() - t0) / steps * 1e3
    out = {"plain_fp64_ms": plain, "filtered_ms": single, "filtered_batched_ms": wall,
           "filtered_batched_device_ms": float(np.mean(dev)), "speedup_filtered": plain / single,
           "speedup_filtered_batched": plain / wall}
    batch.close()
    log(f"1.31M pair: plain FP64 {plain:.2f} ms, filtered {single:.2f} ms, filtered batched {wall:.2f} ms "
        f"(device {out['filtered_batched_device_ms']:.2f})")
    return out


def dropin_section(mg, steps: int) -> dict:
    """The whole boolean where the reference's combine() succeeds: the
    subdivision-6 pair (81,920 triangles each, seed 1, offset (0.5, 0.2, 0.1))
    through the drop-in SolidMesh / SolidBoolean (GPU hot path + host weld,
    retriangulation, face groups), next to the compiled reference's own
    combine() on this host (one core) when it is available."""
    from solidboolean_amd import dropin

    a, b = mg.icosphere_pair(6, 1, (0.5, 0.2, 0.1))

    def run():
        A, B = dropin.SolidMesh(a.vertices, a.triangles), dropin.SolidMesh(b.vertices, b.triangles)
        A.prepare()
        B.prepare()
        sb = dropin.SolidBoolean(A, B)
        ok = sb.combine()
        return ok, sb, (A, B)

    run()
    times = []
    for _ in range(max(3, steps // 2)):
        t0 = time.perf_counter()
        ok, sb, keep = run()
        times.append(time.perf_counter() - t0)
    counts = [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())]
    out = {"workload": "subdivision-6 icosphere pair (81,920 triangles each), seed 1, offset (0.5, 0.2, 0.1)",
           "combine_ok": ok, "face_counts": counts, "ms": float(np.median(times)) * 1e3,
           "stage_seconds": sb.stage_seconds()}
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    try:
        import reflib

        if reflib.available():
            L = reflib.RefLib()
            rt = []
            for _ in range(3):
                t0 = time.perf_counter()
                ra, rb = L.mesh(a.vertices, a.triangles), L.mesh(b.vertices, b.triangles)
                bo = reflib.RefBoolean(ra, rb)
                rt.append(time.perf_counter() - t0)
            out["reference_cpu_ms"] = float(np.median(rt)) * 1e3
            out["reference_face_counts"] = [len(bo.fetch(i)) for i in range(3)]
            out["reference_combine_ok"] = bo.ok
            out["speedup_vs_reference"] = out["reference_cpu_ms"] / out["ms"]
    except ImportError:
        pass
    log(f"drop-in combine (sub-6 pair): {out['ms']:.1f} ms, faces {counts}"
        + (f"; reference {out['reference_cpu_ms']:.1f} ms" if "reference_cpu_ms" in out else ""))
    return out


def dropin_large_section(a, b, reps: int = 3) -> dict:
    """The whole boolean on the 1.31M-triangle pair through the drop-in with
    the robust weld (the reference's combine() fails on this pair, and so
    does the drop-in's default weld, which follows it)."""
    from solidboolean_amd import dropin

    times, sb = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        A, B = dropin.SolidMesh(a.vertices, a.triangles), dropin.SolidMesh(b.vertices, b.triangles)
        A.prepare()
        B.prepare()
        sb = dropin.SolidBoolean(A, B, robust_weld=True)
        ok = sb.combine()
        times.append(time.perf_counter() - t0)
    out = {"workload": "the 1,310,720-triangle pair, robust weld", "combine_ok": ok,
           "ms": float(np.median(times)) * 1e3, "stage_seconds": sb.stage_seconds(),
           "face_counts": [len(sb.fetch_union()), len(sb.fetch_diff()), len(sb.fetch_intersect())] if ok else None,
           "reference_combine_ok": False}
    log(f"drop-in robust combine (1.31M pair): {out['ms']:.1f} ms, ok {ok}")
    return out


def stress_section(ctx, mg, N, steps: int) -> dict:
    """The filter stress cases: per case the candidate count, hits, coplanar
    pairs, the FP32 filter's fallback rate, both GD precisions per launch, and
    the end-to-end step (prepare both meshes, search, filtered GD, classify
    both)."""
    out = {}
    cases = (("torus_4194304_vs_sphere_327680", lambda: mg.torus_sphere_pair(1)),
             ("cube_grids_3145728_x2_near_coplanar", lambda: mg.cube_pair(512)))
    for name, make in cases:
        a, b = make()
        A = ctx.mesh(a.vertices, a.triangles, prepare=False)
        B = ctx.mesh(b.vertices, b.triangles, prepare=False)
        votes = (N.host_array(A.nt), N.host_array(B.nt))  # page-locked: the votes arrive by DMA

        def step():
            A.prepare()
            B.prepare()
            P = A.find_pairs(B)
            P.tri_tri(A, B, N.FP32_FILTERED, stats=False)
            A.classify_pair(B, P, out=votes)
            return P

        step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            P = step()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        st = P.tri_tri(A, B, N.FP32_FILTERED, stats=True)
        f_us = P.bench(A, B, N.FP32_FILTERED, 50) * 1e3
        d_us = P.bench(A, B, N.FP64, 50) * 1e3
        out[name] = {"triangles": [a.nt, b.nt], "candidate_pairs": st["tested"], "hits": st["hits"],
                     "coplanar": st["coplanar"], "fallback_pairs": st["fallback_pairs"],
                     "fallback_rate": st["fallback_pairs"] / max(1, st["tested"]),
                     "filtered_us_per_launch": f_us, "fp64_us_per_launch": d_us, "end_to_end_ms": ms}
        log(f"stress {name}: {st['tested']} pairs, fallback {st['fallback_pairs']}, {ms:.2f} ms/step")
        del P, A, B
        # the whole boolean through the drop-in, default weld (as the
        # reference) and the robust weld chain
        from solidboolean_amd import dropin

        for robust in (False, True):
            t0 = time.perf_counter()
            DA, DB = dropin.SolidMesh(a.vertices, a.triangles), dropin.SolidMesh(b.vertices, b.triangles)
            DA.prepare()
            DB.prepare()
            ok = dropin.SolidBoolean(DA, DB, robust_weld=robust).combine()
            key = "dropin_combine_robust" if robust else "dropin_combine_default"
            out[name][key] = {"ok": ok, "ms": (time.perf_counter() - t0) * 1e3}
            del DA, DB
        log(f"stress {name}: drop-in default {out[name]['dropin_combine_default']}, "
            f"robust {out[name]['dropin_combine_robust']}")
    return out


def batch_section(ctx, mg, N, steps: int, graph: bool = True) -> dict:
    """4,096 independent icosphere pairs (levels 1-6, P ~ 4^-level, radii
    U[0.5, 2], offsets U[-0.8, 0.8]^3 x radius, seeded rotations) as one batch:
    autotuned, then timed as graph replays."""
    t0 = time.perf_counter()
    pairs = mg.batch_pairs(4096, seed=1)
    batch = ctx.batch(pairs)
    log(f"batch: {len(pairs)} pairs, {batch.nt} triangles, built in {time.perf_counter() - t0:.1f}s")
    if not graph:
        batch.set_tuning(use_graph=0)
    batch.run(N.FP32_FILTERED)
    tuned, tuned_ms = batch.autotune(N.FP32_FILTERED, reps=2)
    log(f"batch autotune: {tuned} ({tuned_ms:.3f} ms)")
    res = {}
    for name, prec in (("filtered", N.FP32_FILTERED), ("fp64", N.FP64)):
        batch.run(prec)
        batch.run(prec)
        dev, t0 = [], time.perf_counter()
        for _ in range(steps):
            dev.append(batch.run(prec))
        wall = (time.perf_counter() - t0) / steps * 1e3
        res[name] = (wall, float(np.mean(dev)))
    ncand, nhits = batch.counts()
    wall, dev = res["filtered"]
    # the Guigue-Devillers kernel alone over the batch's 1.15M distinct
    # candidates (the last run's list): its roofline in the plan's units
    gd = {}
    for name, prec in (("filtered", N.FP32_FILTERED), ("fp64", N.FP64)):
        gd[name] = batch.bench_tri_tri(prec, 50)
    gms, gst = gd["filtered"]
    gbytes = gd_bytes(ncand, gst["hits"], gst["fallback_pairs"], 1)
    gtraffic = measured_traffic(("k_gd_filter", "k_gd_list"), None, GD_TRAFFIC_PROFILE)
    roof_gd = {
        "kernel": "k_gd_filter + k_gd_list (FP32-filtered Guigue-Devillers) over the 4,096-pair batch's candidates",
        "bound": "hbm",
        "achieved": gbytes / (gms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": gbytes / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": gtraffic,
        "traffic_frac": gtraffic / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS if gtraffic else None,
        "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/hbm_traffic_pmc_gd_4096_batch.md",
        "bytes_per_launch": gbytes,
        "bytes_units": "per pair 2 x 36 B FP32 coordinates + 8 B ids + 1 B status; per FP64 re-run 2 x 72 B; "
                       "per hit 48 B segment",
        "avg_launch_ms": gms,
        "pairs": ncand,
        "fallback_pairs": gst["fallback_pairs"],
        "fp64_avg_launch_ms": gd["fp64"][0],
        "filtered_speedup_vs_fp64": gd["fp64"][0] / gms,
        "tests_per_s": ncand / (gms * 1e-3),
        "fp64_tests_per_s": ncand / (gd["fp64"][0] * 1e-3),
    }
    log(f"batch GD kernel: filtered {gms * 1e3:.1f} us, fp64 {gd['fp64'][0] * 1e3:.1f} us per launch over {ncand} "
        f"candidates, fallback {gst['fallback_pairs']}")
    out = {
        "workload": "4,096 icosphere pairs, levels 1-6 with P ~ 4^-level, radii U[0.5, 2], offsets U[-0.8, 0.8]^3 "
                    "x radius, seeded rotations (seed 1), " + ("one captured graph per run" if graph else
                                                              "plain launches (--no-graph)"),
        "pairs": len(pairs),
        "triangles": batch.nt,
        "candidate_pairs": ncand,
        "intersecting_pairs": nhits,
        "ms_per_batch": wall,
        "device_ms_per_batch": dev,
        "mesh_pairs_per_s": len(pairs) / (wall * 1e-3),
        "tri_pair_tests_per_s": ncand / (wall * 1e-3),
        "fp64_ms_per_batch": res["fp64"][0],
        "filtered_speedup_vs_fp64": res["fp64"][0] / wall,
        "tuning": tuned,
        "roofline_gd": roof_gd,
    }
    batch.close()
    log(f"batch: {wall:.2f} ms per batch (device {dev:.2f}), fp64 {res['fp64'][0]:.2f} ms")
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4,096-pair batch and the FP64 comparison")
    ap.add_argument("--no-graph", action="store_true", help="run the batch as plain launches (for kernel tracing)")
    ap.add_argument("--no-stress", action="store_true", help="skip the torus / cube filter stress cases")
    ap.add_argument("--gd-only", type=int, default=0, help="profiling aid: only N back-to-back GD launches")
    ap.add_argument("--step-only", type=int, default=0, help="profiling aid: only N steps with default knobs")
    ap.add_argument("--rays-only", type=int, default=0,
                    help="profiling aid: one step, then N launches of the x/y ray kernel and a 1 GiB copy")
    ap.add_argument("--grid-density", type=int, default=0, help="grid density for --step-only / --rays-only")
    ap.add_argument("--batch-only", type=int, default=0,
                    help="profiling aid: the 4,096-pair batch, N runs per precision, N GD launches over its "
                         "candidates per precision, and a 1 GiB copy")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the hot path of one mesh pair "
              f"is measured on one GPU's share; sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    from solidboolean_amd import meshgen as mg
    from solidboolean_amd import native as N

    if args.batch_only:
        ctx = N.Context(0)
        batch = ctx.batch(mg.batch_pairs(4096, seed=1))
        if args.no_graph:
            batch.set_tuning(use_graph=0)
        for name, prec in (("filtered", N.FP32_FILTERED), ("fp64", N.FP64)):
            ms = [batch.run(prec) for _ in range(args.batch_only)]
            gms, gst = batch.bench_tri_tri(prec, args.batch_only)
            log(f"batch-only {name}: {np.median(ms):.3f} ms per run (device), GD {gms * 1e3:.1f} us per launch over "
                f"{gst['tested']} candidates, fallback {gst['fallback_pairs']}")
        cms = ctx.calibrate_copy(1 << 30, 3)
        log(f"copy 1 GiB {cms:.3f} ms")
        batch.close()
        ctx.close()
        return

    t0 = time.perf_counter()
    a, b = mg.large_pair(args.seed, args.level)
    log(f"meshes: {a.nt} + {b.nt} triangles generated in {time.perf_counter() - t0:.1f}s")
    ctx = N.Context(0)
    A = ctx.mesh(a.vertices, a.triangles, prepare=False)
    B = ctx.mesh(b.vertices, b.triangles, prepare=False)
    # the vote arrays are reused step after step, as a pipeline would
    votes = (N.host_array(A.nt), N.host_array(B.nt))  # page-locked: the votes arrive by DMA

    def step(precision=N.FP32_FILTERED):
        A.prepare()
        B.prepare()
        P = A.find_pairs(B)
        P.tri_tri(A, B, precision, stats=False)
        ca, cb = A.classify_pair(B, P, out=votes)
        return P, ca, cb

    if args.gd_only:
        P, _, _ = step()
        ms = P.bench(A, B, N.FP32_FILTERED, args.gd_only)
        ms64 = P.bench(A, B, N.FP64, args.gd_only)
        reps = max(1, (1 << 24) // max(1, len(P)))
        Psat = ctx.pairs(np.tile(P.array(), (reps, 1)))
        sat = Psat.bench(A, B, N.FP32_FILTERED, 5)
        sat64 = Psat.bench(A, B, N.FP64, 5)
        cms = ctx.calibrate_copy(1 << 30, 5)
        log(f"gd-only: {len(P)} pairs, filtered {ms * 1e3:.2f} us, fp64 {ms64 * 1e3:.2f} us per launch; "
            f"saturated {len(Psat)} pairs: filtered {sat:.3f} ms, fp64 {sat64:.3f} ms; copy 1 GiB {cms:.3f} ms")
        del Psat, P
        return

    if args.grid_density:
        ctx.set_tuning(grid_density=args.grid_density)
    if args.rays_only:
        P, _, _ = step()
        for rp in (N.FP64, N.FP32_FILTERED):
            ctx.set_tuning(ray_precision=rp)
            ms, work = A.bench_rays(B, P, 0, args.rays_only)
            log(f"rays-only: grid density {ctx.tuning()['grid_density']}, ray precision {rp}, x/y ray kernel "
                f"{ms * 1e3:.1f} us per launch, {work}")
        cms = ctx.calibrate_copy(1 << 30, 5)
        log(f"copy 1 GiB {cms:.3f} ms")
        return

    if args.step_only:
        for i in range(args.step_only):
            step()
            ctx.synchronize()
            log(f"step {i + 1}/{args.step_only} done")
        cms = ctx.calibrate_copy(1 << 30, 3)
        log(f"copy 1 GiB {cms:.3f} ms")
        return

    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    tuning, tuned_ms = ctx.autotune_pair(A, B, reps=2)
    log(f"autotuned single-pair knobs: {tuning} ({tuned_ms:.2f} ms prepare + search + GD + classify)")
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        P, ca, cb = step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    # stage breakdown from HIP events in a second, separately timed run (the
    # event records are kept out of the wall-clock steps above)
    ctx.stage_ms(reset=True)
    ctx.set_timing(True)
    for i in range(args.steps):
        P, ca, cb = step()
    ctx.synchronize()
    ctx.set_timing(False)
    stages = {k: v / args.steps for k, v in ctx.stage_ms(reset=True).items()}
    log(f"timed {args.steps} steps: {elapsed / args.steps * 1e3:.2f} ms/step; stages {stages}")
    n_pairs = len(P)

    # per-step outputs, checked against the reference's digests when available
    stats = P.tri_tri(A, B, N.FP32_FILTERED, stats=True)
    hp, hs = P.hits()
    parity = None
    gpath = os.path.join(ROOT, "tests", "golden", "large_s1.npz")
    if args.seed == 1 and args.level == 8 and os.path.exists(gpath):
        with np.load(gpath, allow_pickle=False) as g:
            parity = {
                "pairs": sha(P.array().astype("<u4")) == str(g["pair_digest"]),
                "segments": sha(hp.astype("<u4"), np.ascontiguousarray(hs, "<f8")) == str(g["hits_digest"]),
                "classification": sha(ca) == str(g["cls_digest_a"]) and sha(cb) == str(g["cls_digest_b"]),
            }

    # Guigue-Devillers kernel: filtered vs plain FP64, natural and saturated
    iters = 200
    gd_ms = P.bench(A, B, N.FP32_FILTERED, iters)
    gd64_ms = P.bench(A, B, N.FP64, iters)
    ab = P.array()
    reps = max(1, (1 << 24) // max(1, n_pairs))
    Psat = ctx.pairs(np.tile(ab, (reps, 1)))
    sat_ms = Psat.bench(A, B, N.FP32_FILTERED, 20)
    sat64_ms = Psat.bench(A, B, N.FP64, 20)
    n_sat = len(Psat)
    hits = stats["hits"]
    fb = stats["fallback_pairs"]
    tests_per_s = n_pairs / (stages["tri_tri"] * 1e-3) if stages["tri_tri"] > 0 else n_pairs / (gd_ms * 1e-3)

    # roofline of the dominant kernel of the step: the x/y ray kernel of the
    # classification stage, timed alone with HIP events on the context stream
    dom = max(("records", "sort", "bvh", "traverse", "pair_sort", "tri_tri", "compact", "classify", "grid"),
              key=lambda k: stages[k])
    lanes = 2 * A.nt
    density = ctx.tuning()["grid_density"]
    # timed at the per-lane cell limit the committed rocprofv3 summary was
    # taken at (the library default), whatever the autotuner chose for the
    # step, so that the two describe the same launch
    step_cap = ctx.tuning()["ray_cell_cap"]
    rays = {}
    for name, rp in (("fp64", N.FP64), ("filtered", N.FP32_FILTERED)):
        saved = ctx.tuning()
        ctx.set_tuning(**{**saved, "ray_precision": rp, "ray_cell_cap": RAYS_PROFILE_CELL_CAP})
        rays[name] = A.bench_rays(B, P, 0, 20)
        ctx.set_tuning(**saved)
    ray_ms, ray_work = rays["filtered"]
    rbytes = ray_bytes(ray_work, True)
    traffic = measured_traffic(("void k_rays_grid<true, false, true>",), lanes, RAYS_TRAFFIC_PROFILE)
    roof_rays = {
        "kernel": "k_rays_grid, FP32-filtered (x and y rays of mesh A's 1,310,720 triangle centroids into mesh B's "
                  "grid)",
        "bound": "hbm",
        "achieved": rbytes / (ray_ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": rbytes / (ray_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": traffic,
        "traffic_frac": traffic / (ray_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
        "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/hbm_traffic_pmc_rays_1310720_pair.md",
        "bytes_per_launch": rbytes,
        "bytes_units": "25 B per ray (query + result), 8 B per grid entry scanned, 36 B per FP32 triangle test",
        "avg_launch_ms": ray_ms,
        "work_per_launch": ray_work,
        "rays_to_fp64": ray_work["to_fp64"],
        "rays_to_fp64_uncertain": ray_work["to_fp64"] - rays["fp64"][1]["to_fp64"],
        "fallback_rate": (ray_work["to_fp64"] - rays["fp64"][1]["to_fp64"]) / max(1, ray_work["rays"]),
        "fp64_kernel_ms": rays["fp64"][0],
        "fp64_kernel_bytes": ray_bytes(rays["fp64"][1], False),
        "grid_density": density,
        "ray_cell_cap": RAYS_PROFILE_CELL_CAP,
        "step_ray_cell_cap": step_cap,
    }
    log(f"x/y ray kernel: filtered {ray_ms * 1e3:.1f} us, fp64 {rays['fp64'][0] * 1e3:.1f} us per launch, "
        f"{ray_work}, {rbytes / 1e6:.0f} MB algorithmic")
    # the same step with the inputs handed over as host arrays (sbg_mesh_create
    # uploads them), i.e. including the PCIe copies
    pt = []
    for _ in range(3):
        t1 = time.perf_counter()
        A2 = ctx.mesh(a.vertices, a.triangles, prepare=False)
        B2 = ctx.mesh(b.vertices, b.triangles, prepare=False)
        A2.prepare()
        B2.prepare()
        P2 = A2.find_pairs(B2)
        P2.tri_tri(A2, B2, N.FP32_FILTERED, stats=False)
        A2.classify_pair(B2, P2)
        pt.append(time.perf_counter() - t1)
        del P2, A2, B2
    pcie_ms = float(np.median(pt)) * 1e3
    log(f"step with host inputs (uploads included): {pcie_ms:.2f} ms")

    # face groups of mesh A's uncut triangles (sbg_components), host arrays in
    # and out as the call takes them
    blocked = np.zeros(A.nt, np.uint8)
    blocked[np.unique(hp[:, 0])] = 1
    labels = A.components(blocked)
    ct = []
    for _ in range(5):
        t1 = time.perf_counter()
        A.components(blocked)
        ct.append(time.perf_counter() - t1)
    comp = {
        "ms": float(np.median(ct)) * 1e3,
        "triangles": A.nt,
        "blocked": int(blocked.sum()),
        "groups": int(np.unique(labels[labels != 0xFFFFFFFF]).size),
    }
    log(f"components: {comp}")
    cal_bytes = 1 << 30
    copy_ms = ctx.calibrate_copy(cal_bytes, 10)
    copy_gbs = 2 * cal_bytes / (copy_ms * 1e-3) / 1e9

    line = {
        "metric": "tri-pair tests/s + end-to-end boolean ms (1.31M-tri pair), filtered vs plain FP64",
        "value": tests_per_s,
        "unit": "tri-pair tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (fp32 sign filter)",
        "data": "synthetic (seeded icospheres)",
        "config": {
            "workload": "icosphere pair, subdivision 8 (1,310,720 triangles each), radial noise 1e-3, "
                        "B rotated (seed 1) and offset by (0.4, 0.25, 0.1)",
            "candidate_pairs": n_pairs,
            "intersecting_pairs": hits,
            "parallelism": "single GPU",
            "tuning": tuning,
        },
        "hot_path_step_ms": elapsed / args.steps * 1e3,
        "hot_path_step_ms_with_uploads": pcie_ms,
        "stage_ms": stages,
        "tri_tri": {
            "filtered_us_per_launch": gd_ms * 1e3,
            "fp64_us_per_launch": gd64_ms * 1e3,
            "filtered_tests_per_s": n_pairs / (gd_ms * 1e-3),
            "fp64_tests_per_s": n_pairs / (gd64_ms * 1e-3),
            "saturated_tests": n_sat,
            "saturated_filtered_tests_per_s": n_sat / (sat_ms * 1e-3),
            "saturated_fp64_tests_per_s": n_sat / (sat64_ms * 1e-3),
            "filtered_speedup_vs_fp64": gd64_ms / gd_ms,
            "saturated_speedup_vs_fp64": sat64_ms / sat_ms,
            "fallback_pairs": fb,
            "fallback_rate": fb / max(1, n_pairs),
            "exits": {k: stats[k] for k in ("exit_plane1", "exit_plane2", "exit_interval", "hits", "coplanar")},
        },
        "dominant_stage": dom,
        "roofline": roof_rays,
        "components": comp,
        "copy_kernel_gbs": copy_gbs,
        "parity": parity,
    }
    if not args.no_stress:
        line["stress"] = stress_section(ctx, mg, N, args.steps)
        line["dropin_combine"] = dropin_section(mg, args.steps)
        line["dropin_combine_1310720_robust"] = dropin_large_section(a, b)
        # the whole boolean (hot path plus host weld, re-triangulation and
        # face groups, meshes handed over as host arrays) beside the hot path
        line["whole_boolean_ms"] = line["dropin_combine_1310720_robust"]["ms"]
    if not args.no_batch:
        line["fp64_vs_filtered_batched"] = fp64_vs_batched(ctx, a, b, step, N, args.steps, not args.no_graph)
        line["batch"] = batch_section(ctx, mg, N, args.steps, not args.no_graph)
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(a, b, None)
    print(json.dumps(line), flush=True)
    del Psat, A, B
    ctx.close()


if __name__ == "__main__":
    main()
