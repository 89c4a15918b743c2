"""Workgroup-size tuning of the batched multigrid kernels (include/mpr_hip.h mpr_autotune).

The colour sweep, residual and FP64->FP32 defect kernels of every V-cycle (solo and batched)
launch with a workgroup size looked up per (kernel, precision, grid size class). autotune()
times the candidates on the device and keeps the fastest for the rest of the process. The
per-node arithmetic does not depend on the workgroup size, so results stay bit-identical.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib, load_library

KERNELS = ("relax", "residual", "defect")
CANDIDATES = (64, 128, 256, 512, 1024)
SIZE_CLASSES = 11


def autotune(max_depth: int = 8, reps: int = 20) -> dict:
    """Times every candidate workgroup size for grid classes 5..max_depth and keeps the
    fastest. Returns {(kernel, precision, class): block} for the tuned entries."""
    table = np.zeros((len(KERNELS), 2, SIZE_CLASSES), np.int32)
    check(lib().mpr_autotune(int(max_depth), int(reps),
                             table.ctypes.data_as(ctypes.c_void_p)), "mpr_autotune")
    out = {}
    for k, name in enumerate(KERNELS):
        for t, prec in enumerate(("fp32", "fp64")):
            if name == "defect" and prec == "fp32":
                continue
            for s in range(5, max_depth + 1):
                out[(name, prec, s)] = int(table[k, t, s])
    return out


def tuned_block(kernel: str, precision: str, size_class: int) -> int:
    v = load_library().mpr_tune_get(KERNELS.index(kernel), int(precision == "fp64"), int(size_class))
    check(-1 if v < 0 else 0, "mpr_tune_get")
    return v


def set_tuned_block(kernel: str, precision: str, size_class: int, block: int) -> None:
    check(load_library().mpr_tune_set(KERNELS.index(kernel), int(precision == "fp64"), int(size_class),
                             int(block)), "mpr_tune_set")
