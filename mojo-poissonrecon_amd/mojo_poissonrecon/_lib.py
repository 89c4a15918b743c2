"""ctypes loader for the HIP library libmpr_hip.so (reference: python/mojo_poissonrecon/_lib.py).

Mirrors the reference module: ``lib()``, ``build()``, ``BuildError`` and the buffer
conversion helpers ``f64``/``i64``/``addr`` with the same rules and messages
(_lib.py:93-145). ``_SIGNATURES`` holds the reference's eleven kernel symbols with their
argument lists unchanged (_lib.py:20-32: int64 addresses and doubles), plus the
device-resident pipeline of include/mpr_hip.h.

There is no CPU fallback: ``lib()`` raises if the library is not built or no HIP device
is present.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _build

LIB = _build.LIB

I = ctypes.c_int64
F = ctypes.c_double
P = ctypes.c_void_p
S = ctypes.c_int32


class mpr_params(ctypes.Structure):
    _fields_ = [
        ("point_weight", ctypes.c_double),
        ("samples_per_node", ctypes.c_double),
        ("scale", ctypes.c_double),
        ("cycles", ctypes.c_int32),
        ("confidence", ctypes.c_int32),
        ("pre_smooth", ctypes.c_int32),
        ("post_smooth", ctypes.c_int32),
        ("stages", ctypes.c_int32),
        ("inner_cycles", ctypes.c_int32),
    ]


class mpr_result(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_int64),
        ("vertex_count", ctypes.c_int64),
        ("face_count", ctypes.c_int64),
        ("iso_value", ctypes.c_double),
        ("residual_norm", ctypes.c_double),
        ("origin", ctypes.c_double * 3),
        ("width", ctypes.c_double),
        ("valid_rows", ctypes.c_int64),
    ]


PRECISION_FP64 = 0
PRECISION_MIXED = 1
STAGE_NORMALIZE = 1
STAGE_SOLVE = 2
STAGE_EXTRACT = 4
STAGE_GRIDS = 8
STAGE_WORLD = 16
STAGE_FILTER = 32

# the reference's kernel table (_lib.py:20-32), unchanged
_SIGNATURES = {
    "mpr_parallel_init": ([], I),
    "mpr_octree_codes_f64": ([I, I, I, I], None),
    "mpr_splat_f64": ([I] * 10 + [F], I),
    "mpr_assemble_f64": ([I, I, I, I, I, I, F, F], None),
    "mpr_relax_f64": ([I, I, I, I, F, I], None),
    "mpr_residual_f64": ([I, I, I, I, I, F], F),
    "mpr_restrict_f64": ([I, I, I], None),
    "mpr_prolong_add_f64": ([I, I, I], None),
    "mpr_iso_value_f64": ([I, I, I, I, I, I], F),
    "mpr_dual_vertices_f64": ([I, I, I, F, I, I, I, I], I),
    "mpr_dual_faces_f64": ([I, I, F, I, I, I], I),
}

_PP = ctypes.POINTER(P)
_IP = ctypes.POINTER(ctypes.c_int64)

# device-resident pipeline (include/mpr_hip.h)
_PIPELINE = {
    "mpr_last_error": ([], ctypes.c_char_p),
    "mpr_device_info": ([ctypes.c_char_p, I, _IP, _IP], ctypes.c_int),
    "mpr_device_memory": ([_IP, _IP], ctypes.c_int),
    "mpr_job_footprint": ([S, I, S], I),
    "mpr_job_create": ([S, I, S, _PP], ctypes.c_int),
    "mpr_job_destroy": ([P], ctypes.c_int),
    "mpr_job_set_points": ([P, P, P, I], ctypes.c_int),
    "mpr_job_run": ([P, ctypes.POINTER(mpr_params)], ctypes.c_int),
    "mpr_job_wait": ([P, ctypes.POINTER(mpr_result)], ctypes.c_int),
    "mpr_job_query": ([P], ctypes.c_int),
    "mpr_job_fetch_mesh": ([P, P, P, P], ctypes.c_int),
    "mpr_job_set_host_mesh": ([P, S], ctypes.c_int),
    "mpr_job_host_mesh": ([P, _PP, _PP, _PP], ctypes.c_int),
    "mpr_job_fetch_grids": ([P, P, P, P, P], ctypes.c_int),
    "mpr_job_build_octree": ([P, S, _IP], ctypes.c_int),
    "mpr_job_fetch_octree": ([P, P, P, P, P, P, P], ctypes.c_int),
    "mpr_job_capture": ([P, ctypes.POINTER(mpr_params)], ctypes.c_int),
    "mpr_job_launch": ([P], ctypes.c_int),
    "mpr_job_device_points": ([P, _PP, _PP], ctypes.c_int),
    "mpr_job_set_count": ([P, I], ctypes.c_int),
    "mpr_job_fetch_filtered": ([P, _IP, P, P], ctypes.c_int),
    "mpr_job_time_kernel": ([P, S, S, ctypes.POINTER(F), ctypes.POINTER(F)], ctypes.c_int),
    "mpr_autotune": ([S, S, P, P, P], ctypes.c_int),
    "mpr_tune_get": ([S, S, S], ctypes.c_int),
    "mpr_tune_set": ([S, S, S, S], ctypes.c_int),
    "mpr_tune_get_groups": ([S, S, S], ctypes.c_int),
    "mpr_tune_set_groups": ([S, S, S, S], ctypes.c_int),
    "mpr_tune_get_planes": ([S, S, S], ctypes.c_int),
    "mpr_tune_set_planes": ([S, S, S, S], ctypes.c_int),
    "mpr_tune_get_march": ([], ctypes.c_int),
    "mpr_tune_set_march": ([S], ctypes.c_int),
    "mpr_fast_division_check": ([S, _IP], ctypes.c_int),
    "mpr_batch_create": ([P, S, _PP], ctypes.c_int),
    "mpr_batch_destroy": ([P], ctypes.c_int),
    "mpr_batch_run": ([P, ctypes.POINTER(mpr_params)], ctypes.c_int),
    "mpr_batch_capture": ([P, ctypes.POINTER(mpr_params)], ctypes.c_int),
    "mpr_batch_time_kernel": ([P, S, S, ctypes.POINTER(F), ctypes.POINTER(F)], ctypes.c_int),
    "mpr_batch_launch": ([P], ctypes.c_int),
    "mpr_batch_query": ([P], ctypes.c_int),
    "mpr_batch_wait": ([P, ctypes.POINTER(mpr_result)], ctypes.c_int),
    "mpr_relax_f32": ([P, P, P, I, F, I], ctypes.c_int),
    "mpr_residual_f32": ([P, P, P, P, I, F], ctypes.c_int),
    "mpr_restrict_f32": ([P, P, I], ctypes.c_int),
    "mpr_prolong_add_f32": ([P, P, I], ctypes.c_int),
    "mpr_vcycle_transfer_f32": ([P, P, P, P, P, I, F], ctypes.c_int),
    "mpr_prolong_c1_f32": ([P, P, I], ctypes.c_int),
    "mpr_extract": ([P, P, I, F, _PP, _IP, _IP], ctypes.c_int),
    "mpr_mesh_fetch": ([P, P, P, P], ctypes.c_int),
    "mpr_mesh_destroy": ([P], ctypes.c_int),
    "mpr_iso_value_f64_noalloc": ([P, P, P, I, I, I, P, P], ctypes.c_int),
    "mpr_mesh_distance": ([P, I, P, I, P, I, P, I, P], ctypes.c_int),
    "mpr_oriented_points_from_mesh": ([P, I, P, I, P, P, _IP], ctypes.c_int),
}


class BuildError(RuntimeError):
    pass


class DeviceError(RuntimeError):
    """A HIP call failed or no GPU is available."""


def build(force: bool = False) -> str:
    """Compile libmpr_hip.so for gfx950 (reference: _lib.py:55-71 rebuilds when stale)."""
    try:
        return _build.build(force=force or not _build.up_to_date())
    except _build.BuildError as error:
        raise BuildError(str(error)) from error


_loaded: ctypes.CDLL | None = None


def load_library() -> ctypes.CDLL:
    """Load the shared library and declare signatures, without touching the GPU."""
    global _loaded
    if _loaded is None:
        if not os.path.exists(LIB):
            raise BuildError(f"{LIB} is not built; call mojo_poissonrecon.build() first")
        handle = ctypes.CDLL(LIB)
        for table in (_SIGNATURES, _PIPELINE):
            for name, (argtypes, restype) in table.items():
                fn = getattr(handle, name)
                fn.argtypes = argtypes
                fn.restype = restype
        _loaded = handle
    return _loaded


_device_ready = False


def lib() -> ctypes.CDLL:
    """The library with an initialised GPU (reference: _lib.py:77-90)."""
    global _device_ready
    handle = load_library()
    if not _device_ready:
        if handle.mpr_parallel_init() != 0:
            raise DeviceError("no HIP device available: " + last_error())
        _device_ready = True
    return handle


def last_error() -> str:
    message = load_library().mpr_last_error()
    return message.decode() if message else ""


def check(status: int, what: str) -> None:
    if status != 0:
        raise DeviceError(f"{what} failed ({status}): {last_error()}")


def f64(values, *, copy: bool = False) -> np.ndarray:
    source = np.asarray(values)
    if np.issubdtype(source.dtype, np.complexfloating):
        raise TypeError("complex values cannot be converted to Float64")
    if not (np.issubdtype(source.dtype, np.number) or np.issubdtype(source.dtype, np.bool_)):
        raise TypeError("values must be real numbers")
    if np.issubdtype(source.dtype, np.floating) and source.dtype.itemsize > 8:
        raise TypeError("extended-precision floats cannot be narrowed to Float64")
    if np.issubdtype(source.dtype, np.integer):
        limit = 1 << 53
        if source.size and (np.any(source > limit) or np.any(source < -limit)):
            raise ValueError("integer value cannot be represented exactly as Float64")
    if copy:
        return np.array(values, dtype=np.float64, order="C", copy=True)
    return np.ascontiguousarray(values, dtype=np.float64)


def i64(values, *, copy: bool = False) -> np.ndarray:
    source = np.asarray(values)
    if np.issubdtype(source.dtype, np.complexfloating):
        raise TypeError("complex values cannot be converted to Int64")
    if np.issubdtype(source.dtype, np.floating):
        if source.size and (not np.isfinite(source).all() or np.any(source != np.trunc(source))):
            raise ValueError("Int64 values must be finite integers")
    elif not (np.issubdtype(source.dtype, np.integer) or np.issubdtype(source.dtype, np.bool_)):
        raise TypeError("values cannot be converted safely to Int64")
    info = np.iinfo(np.int64)
    if source.size and (np.any(source < info.min) or np.any(source > info.max)):
        raise OverflowError("value is outside the Int64 range")
    if copy:
        return np.array(values, dtype=np.int64, order="C", copy=True)
    return np.ascontiguousarray(values, dtype=np.int64)


def addr(values: np.ndarray) -> int:
    if not isinstance(values, np.ndarray):
        raise TypeError("FFI buffers must be NumPy arrays")
    if values.dtype not in (np.dtype(np.float64), np.dtype(np.int64)):
        raise TypeError("FFI buffers must have native Float64 or Int64 dtype")
    if not values.flags.c_contiguous or not values.flags.aligned:
        raise ValueError("FFI buffers must be aligned and C-contiguous")
    address = int(values.ctypes.data)
    if address == 0:
        raise ValueError("FFI buffers must have a non-null address")
    return address


def ptr(values: np.ndarray | None) -> int | None:
    """Address of an (possibly empty) C-contiguous array, or None."""
    if values is None:
        return None
    assert values.flags.c_contiguous
    return int(values.ctypes.data) if values.size else None
