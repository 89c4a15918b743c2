"""Launch-shape tuning of the batched multigrid kernels (include/mpr_hip.h mpr_autotune).

The colour sweep, residual and FP64->FP32 defect kernels of every V-cycle (solo and batched)
launch with a workgroup size looked up per (kernel, precision, grid size class); the colour
sweep also with 1 or 2 slot groups per lane, or in z-column form with one slot group in each
of 2 or 4 consecutive planes per lane. autotune() times the candidates on the device and
keeps the fastest for the rest of the process. The per-node arithmetic does not depend on the launch shape, so results stay
bit-identical.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib, load_library

KERNELS = ("relax", "residual", "defect")
CANDIDATES = (64, 128, 256, 512, 1024)
GROUP_CANDIDATES = (1, 2)
PLANE_CANDIDATES = (1, 2, 4)
SIZE_CLASSES = 11


def autotune(max_depth: int = 8, reps: int = 20) -> dict:
    """Times every candidate launch shape for grid classes 5..max_depth and keeps the
    fastest. Returns {(kernel, precision, class): (block, groups per lane, planes per lane)}."""
    blocks = np.zeros((len(KERNELS), 2, SIZE_CLASSES), np.int32)
    groups = np.zeros_like(blocks)
    planes = np.zeros_like(blocks)
    check(lib().mpr_autotune(int(max_depth), int(reps), blocks.ctypes.data_as(ctypes.c_void_p),
                             groups.ctypes.data_as(ctypes.c_void_p),
                             planes.ctypes.data_as(ctypes.c_void_p)), "mpr_autotune")
    out = {}
    for k, name in enumerate(KERNELS):
        for t, prec in enumerate(("fp32", "fp64")):
            if name == "defect" and prec == "fp32":
                continue
            for s in range(5, max_depth + 1):
                out[(name, prec, s)] = (int(blocks[k, t, s]), int(groups[k, t, s]),
                                        int(planes[k, t, s]))
    return out


def _key(kernel: str, precision: str, size_class: int):
    return KERNELS.index(kernel), int(precision == "fp64"), int(size_class)


def tuned_block(kernel: str, precision: str, size_class: int) -> int:
    v = load_library().mpr_tune_get(*_key(kernel, precision, size_class))
    check(-1 if v < 0 else 0, "mpr_tune_get")
    return v


def set_tuned_block(kernel: str, precision: str, size_class: int, block: int) -> None:
    check(load_library().mpr_tune_set(*_key(kernel, precision, size_class), int(block)),
          "mpr_tune_set")


def tuned_groups(kernel: str, precision: str, size_class: int) -> int:
    v = load_library().mpr_tune_get_groups(*_key(kernel, precision, size_class))
    check(-1 if v < 0 else 0, "mpr_tune_get_groups")
    return v


def set_tuned_groups(kernel: str, precision: str, size_class: int, groups: int) -> None:
    check(load_library().mpr_tune_set_groups(*_key(kernel, precision, size_class), int(groups)),
          "mpr_tune_set_groups")


def tuned_planes(kernel: str, precision: str, size_class: int) -> int:
    v = load_library().mpr_tune_get_planes(*_key(kernel, precision, size_class))
    check(-1 if v < 0 else 0, "mpr_tune_get_planes")
    return v


def set_tuned_planes(kernel: str, precision: str, size_class: int, planes: int) -> None:
    check(load_library().mpr_tune_set_planes(*_key(kernel, precision, size_class), int(planes)),
          "mpr_tune_set_planes")


def march_iterations() -> int:
    """Red-black iterations of each smoothing phase of the mixed-precision V-cycle's FP32
    levels fused into one pass over the level: 0, 1 or 2 (the default). Results are
    bit-identical for every setting."""
    v = load_library().mpr_tune_get_march()
    check(-1 if v < 0 else 0, "mpr_tune_get_march")
    return v


def set_march_iterations(iterations: int) -> None:
    """Sets the fused iterations (0, 1 or 2) for jobs and batches created afterwards."""
    check(load_library().mpr_tune_set_march(int(iterations)), "mpr_tune_set_march")


def fast_division_mismatches(n: int) -> int:
    """Finite float numerators (all 2^32 bit patterns) for which the fused smoothing's
    division by a node's unscreened diagonal deg (n-1)^2, deg = 3..6, differs from the IEEE
    quotient (0: exact everywhere)."""
    out = ctypes.c_int64()
    check(load_library().mpr_fast_division_check(int(n), ctypes.byref(out)),
          "mpr_fast_division_check")
    return out.value
