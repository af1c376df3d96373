"""Builds libdigholo.so in-tree (HIP kernels for gfx950 + the C API host code).

    python -m digholo_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(os.path.dirname(PKG), "_build")
LIB = os.path.join(PKG, "libdigholo.so")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# hipcc flags for device code: gfx950 only, packed-FP32 instructions disabled
DEVICE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# host-only C++ for the API: no FMA contraction, so float32 geometry and the
# Fourier mask round exactly as the reference's scalar/AVX code does
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-result"]

KERNEL_SOURCES = ["dh_kernels.hip", "dh_fused.hip", "dh_overlap.hip", "dh_metrics.hip", "dh_analysis.hip",
                  "dh_align.hip", "dh_grad.hip", "dh_basis.hip", "dh_refcal.hip", "dh_noalloc.hip"]
HOST_SOURCES = ["dh_api.cpp", "dh_eigen.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True, jobs: int = 0) -> str:
    """Compiles every source (in parallel, `jobs` at a time; 0: one per CPU up
    to 8) and links libdigholo.so."""
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src in KERNEL_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        # the FP64 eigen-solve behind mode-dependent loss runs on the host cores
        omp = ["-fopenmp"] if src == "dh_eigen.cpp" else []
        cmds.append(["g++", *HOST_FLAGS, *omp, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the C++-linkage names of every entry point (what a caller of the
    # reference header links against), generated from include/digHolo.h
    from digholo_amd.gen_cxx_aliases import generate
    alias_src = os.path.join(BUILD, "dh_cxx_aliases.cpp")
    with open(alias_src, "w") as f:
        f.write(generate())
    obj = alias_src + ".o"
    cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-c", alias_src, "-o", obj])
    objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-lgomp", "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
