"""Compiles the HIP sources under csrc/ into libmpr_hip.so for gfx950 with hipcc.

The library is built in-tree (next to this file) so it travels with the source tree.
Flags that matter for results: ``-ffp-contract=off`` keeps every multiply and add
separately rounded, as the reference computes them; ``-munsafe-fp-atomics`` lets FP64
atomic adds use the hardware instruction on coarse-grained HBM allocations.
"""

from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                       # mojo-poissonrecon_amd/
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(os.path.dirname(ROOT), "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libmpr_hip.so")
ARCH = os.environ.get("MPR_OFFLOAD_ARCH", "gfx950")

SOURCES = ["grid_kernels.hip", "batch_kernels.hip", "vcycle.hip", "point_kernels.hip",
           "extract_kernels.hip", "solver.hip",
           "capi_compat.hip", "tune.hip", "noalloc.hip",
           "metric_kernels.hip", "mesh_kernels.hip", "march.hip"]

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


class BuildError(RuntimeError):
    pass


def _hipcc() -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        raise BuildError("hipcc not found (ROCm is required to build libmpr_hip.so)")
    return hipcc


def _inputs() -> list[str]:
    files = [os.path.join(CSRC, f) for f in os.listdir(CSRC)]
    files.append(os.path.join(INCLUDE, "mpr_hip.h"))
    return files


def up_to_date() -> bool:
    if not os.path.exists(LIB):
        return False
    stamp = os.path.getmtime(LIB)
    return all(os.path.getmtime(f) <= stamp for f in _inputs())


def build(force: bool = True, jobs: int = 5) -> str:
    """Compile every source (in parallel) and link the shared library."""
    if not force and up_to_date():
        return LIB
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objects = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objects.append(obj)
        cmd = [hipcc, *FLAGS, "-I", CSRC, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
    errors = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode:
            errors.append(f"{src}:\n{err or out}")
    if errors:
        raise BuildError("\n".join(errors)[:8000])
    tmp = LIB + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objects]
    proc = subprocess.run(link, capture_output=True, text=True)
    if proc.returncode:
        raise BuildError((proc.stderr or proc.stdout)[:8000])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
