"""Builds the in-tree HIP library libgraph_render_fa2.so for gfx950 with hipcc.

The library is the product: C-ABI entry points (include/graph_render_fa2.h), host code and
hand-written HIP kernels, one shared object next to this file under _lib/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libgraph_render_fa2.so")
SOURCES = ["fa2_session.cpp", "fa2_kernels.hip", "sort_scan.hip", "pm.hip", "adjoint.hip",
           "snapshot.cpp", "noalloc.hip", "d3.hip", "forcesim.hip", "pm_d3.hip"]
ARCH = os.environ.get("GR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: FP64 kernels must round as the reference's Rust does (no FMA fusion).
# -packed-fp32-ops off: scalar FP32 only (packed forms are avoided on this hardware).
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *[f for f in FLAGS if f != "-shared"], "-c", "-x", "hip",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        for cmd, res in zip(cmds, pool.map(
                lambda c: subprocess.run(c, capture_output=True, text=True), cmds)):
            sys.stderr.write(res.stderr)
            if res.returncode != 0:
                raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
