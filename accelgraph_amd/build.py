"""Build recipe for the in-tree shared libraries.

    python -m accelgraph_amd.build          # libaccelgraph.so (HIP kernels + C host layer)

Products (git-ignored, shipped with the tree):
  accelgraph_amd/libaccelgraph.so   C-ABI of include/accelgraph.h plus the drop-in entry points
                                    of include/pageRank.h, include/SPMV.h, include/graphCSR.h

Device code is compiled for gfx950 only, without FMA contraction (the reference's x86-64 host
arithmetic has none, and the fixed-point paths must match it bit for bit) and with packed-FP32
instruction generation disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "accelgraph_amd")
INCLUDE = os.path.join(ROOT, "include")
HOST_DIR = os.path.join(PKG, "host")
CSRC_DIR = os.path.join(PKG, "csrc")
BUILD_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libaccelgraph.so")

HOST_SOURCES = ["mt19937.c", "graph_csr.c", "graph_io.c", "plan.c", "accel_context.c", "pageRank.c", "SPMV.c"]
HIP_SOURCES = ["ag_runtime.hip", "ag_pagerank.hip", "ag_spmv.hip", "ag_batch.hip", "ag_ppr.hip", "ag_tune.hip", "ag_mlp.hip", "ag_build.hip"]

ARCH = "gfx950"

CFLAGS = ["-O3", "-m64", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
          "-std=gnu11", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{HOST_DIR}"]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-value", f"-I{INCLUDE}", f"-I{HOST_DIR}",
            f"-I{CSRC_DIR}"]


def _hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}")
    if res.stdout.strip():
        # surface compiler warnings without failing the build
        lines = [ln for ln in res.stdout.splitlines()
                 if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")


def _libgomp() -> str:
    out = subprocess.run(["gcc", "-print-file-name=libgomp.so"], stdout=subprocess.PIPE,
                         text=True, check=True).stdout.strip()
    return out


def build(verbose: bool = False) -> str:
    """Compile every source and link libaccelgraph.so; returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        objs.append(obj)
        jobs.append(["gcc", *CFLAGS, "-c", os.path.join(HOST_DIR, src), "-o", obj])
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIPFLAGS, "-c", os.path.join(CSRC_DIR, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, _libgomp(),
          "-Wl,--no-undefined", "-lpthread", "-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
