"""Builds the native library ``riptide_amd/libriptide_amd.so`` in-tree.

HIP sources are compiled for gfx950 with hipcc; the search planner is compiled
with g++ and the reference's floating-point flags (setup.py:13-20,34) so that
its float64 plan arithmetic is bit-identical to the reference's.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libriptide_amd.so")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["ffa_kernels.hip", "ffa_records.hip", "series_kernels.hip", "batch.hip",
               "adjoint.hip", "api.hip", "fir.hip", "peaks.hip", "grad_batch.hip"]
GCC_SOURCES = ["plan.cpp", "ffa_plan.cpp"]
# ffa_variant.hip is compiled once per FFA kernel variant (kFfaVariants in
# csrc/tasks.h), each holding the merge code of its own row lengths
FFA_VARIANT_SOURCE = "ffa_variant.hip"
FFA_VARIANTS = 10

ARCH = os.environ.get("RIPTIDE_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 instructions stay off: some of their forms have given wrong
    # results on this hardware without raising an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no implicit fused multiply-adds: the float-float arithmetic (two-sum,
    # product error terms, remainders) needs every product rounded where it
    # is written; the fused multiply-adds the reference build makes are
    # written out with __fmaf_rn where they matter for parity
    "-ffp-contract=off",
    # no signalling-NaN quieting in min/max (the reference is built with
    # -ffinite-math-only): fmaxf compiles to one v_max_f32
    "-mno-amdgpu-ieee", "-fno-honor-nans",
    "-Wall", "-Wno-unused-function",
]
GCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3",
    "-fassociative-math", "-fno-math-errno", "-ffinite-math-only",
    "-fno-rounding-math", "-fno-signed-zeros", "-fno-trapping-math",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def _jobs():
    try:
        return max(1, min(os.cpu_count() or 1, int(os.environ.get("MAX_JOBS", "8"))))
    except ValueError:
        return 1


def build(verbose=False):
    """Compiles every source (in parallel) and links the shared library
    (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    for v in range(FFA_VARIANTS):
        obj = os.path.join(BUILD, f"ffa_variant_{v}.o")
        cmds.append([HIPCC, *HIP_FLAGS, f"-DRT_FFA_VARIANT={v}", "-c",
                     os.path.join(CSRC, FFA_VARIANT_SOURCE), "-o", obj])
    for src in GCC_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["g++", *GCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    # the slowest compiles (the kernel variants) first
    cmds.sort(key=lambda c: 0 if FFA_VARIANT_SOURCE in c[-3] else 1)
    with ThreadPoolExecutor(_jobs()) as pool:
        list(pool.map(_run, cmds))
    objs = [c[-1] for c in cmds]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
