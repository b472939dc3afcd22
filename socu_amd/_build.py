"""Build libsocu_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m socu_amd._build`` (or ``__graft_entry__.build()``).
The library lands next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsocu_hip.so")

SOURCES = ["nd_blocked.hip", "nd_large.hip", "nd_refine.hip", "nd_batched.hip", "nd_adjoint.hip", "nd_sweep.hip",
           "capi.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # No packed-FP32 instructions anywhere in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def _run(cmd, what):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) {what}")


def build_library(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every source to an object in parallel, then link the shared
    library; the result replaces OUT atomically."""
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    cmds = []
    objs = []
    for s in SOURCES:
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append(([hipcc(), *HIPCC_FLAGS, "-c", "-o", obj, os.path.join(CSRC, s)], f"compiling {s}"))
    if verbose:
        for cmd, _ in cmds:
            print(" ".join(cmd), flush=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, w) for c, w in cmds]:
            f.result()
    tmp = OUT + ".tmp"
    link = [hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    _run(link, f"linking {OUT}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
