"""Builds gale_amd/libgale_stokes.so for gfx950 with hipcc (in-tree).

Run as ``python -m gale_amd.build_lib`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgale_stokes.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["gs_kernels.hip", "gs_tile.hip", "gs_mg.hip", "gs_net.hip", "gs_learned.hip", "gs_batch.hip", "gs_weights.hip", "gs_api.hip", "gs_generate.cpp", "gs_checkpoint.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int = 2) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError(f"hipcc failed on {src}")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    # rocSOLVER (with rocBLAS) factors the dense coarsest multigrid operator
    # at setup when it is too large for the host inverse
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lrocsolver", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
