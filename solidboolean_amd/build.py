"""Builds the native libraries in-tree.

* solidboolean_amd/lib/libsbgpu.so - the HIP kernels + C ABI (include/sbgpu.h),
  compiled for gfx950 with hipcc.
* solidboolean_amd/lib/libsolidboolean.so - the host-side drop-in
  SolidMesh / SolidBoolean C++ API linked against libsbgpu.so (when its
  sources exist).
* solidboolean_amd/lib/sb_driver - the reference's OBJ test driver on the
  drop-in API (solidboolean_amd/driver/sb_driver.cpp).

Every FP64 translation unit is compiled with -ffp-contract=off so that no
a*b+c is fused (the reference's results depend on separate roundings), and
all HIP code with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(ROOT, "_build", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["k_prepare.hip", "k_sort.hip", "k_lbvh.hip", "k_traverse.hip", "k_tritri.hip", "k_classify.hip",
               "k_batch.hip", "k_adjoint.hip", "k_components.hip", "sbg_api.hip", "sbg_batch.hip"]
ARCH = os.environ.get("SBG_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", f"-I{INCLUDE}",
             f"-I{HOST}"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out)


def build_gpu(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cc = hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    out = os.path.join(LIBDIR, "libsbgpu.so")
    tmp = out + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


def host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))


def build_host(verbose: bool = False) -> str | None:
    srcs = host_sources()
    if not srcs:
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libsolidboolean.so")
    tmp = out + ".tmp"
    _run(["g++", *CXX_FLAGS, "-shared", "-o", tmp, *srcs, f"-L{LIBDIR}", "-lsbgpu", "-Wl,-rpath,$ORIGIN",
          "-pthread"], verbose)
    os.replace(tmp, out)
    driver = os.path.join(HERE, "driver", "sb_driver.cpp")
    if os.path.exists(driver):
        exe = os.path.join(LIBDIR, "sb_driver")
        _run(["g++", *[f for f in CXX_FLAGS if f != "-fPIC"], "-o", exe + ".tmp", driver, f"-L{LIBDIR}",
              "-lsolidboolean", "-lsbgpu", "-Wl,-rpath,$ORIGIN", "-pthread"], verbose)
        os.replace(exe + ".tmp", exe)
    return out


def build_all(verbose: bool = False):
    build_gpu(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", os.listdir(LIBDIR))
