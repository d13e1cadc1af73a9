"""Builds bsplineinterpolation_amd/libintp_amd.so for gfx950 (in-tree).

    python -m bsplineinterpolation_amd.build

hipcc compiles the kernels and the C ABI; the host-side setup (knots,
collocation matrices, LU) is compiled with g++ and -ffp-contract=off so its
factors round exactly like the reference's.  Every HIP source is built with
packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libintp_amd.so")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("INTP_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["abi.hip", "batch.hip", "bin.hip", "group.hip", "proxy.hip", "refine.hip", "serve.hip", "solve.hip"] + [
    "batch_eval_d%d.hip" % d for d in (1, 2, 3)] + [
    "eval_%s_d%d.hip" % (t, d) for t in ("f64", "f32") for d in (1, 2, 3, 4, 5, 6)] + [
    "eval_cells_d%d.hip" % d for d in (1, 2, 3)]
CXX_SOURCES = ["setup.cpp", "host_eval.cpp"]

HIPCC_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-ffp-contract=off",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        jobs.append([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
    for src in CXX_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        objs.append(o)
        jobs.append([os.environ.get("CXX", "g++")] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd, _ in zip(jobs, ex.map(_run, jobs)):
            if verbose:
                print(" ".join(cmd))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
