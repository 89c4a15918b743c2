"""Screened Poisson reconstruction on one MI355X (reference: python/mojo_poissonrecon/core.py).

Public names, keyword defaults, dataclass fields, validation order and error messages
follow the reference module. The work runs on the GPU through the device-resident
pipeline of libmpr_hip.so: one upload of the oriented points, one device pass
(normalise, splat, assemble, multigrid V-cycles, iso value, dual-grid extraction) and one
download of the results, instead of one native call per kernel.

Extensions over the reference, all keyword-only with reference-compatible defaults:
``precision`` ("fp64" as the reference, or "mixed": float32 V-cycles inside a float64
defect-correction loop) and depths up to 10 (the reference stops at 8).
"""

from __future__ import annotations

import ctypes
import threading
from contextlib import contextmanager
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import addr, check, f64, i64, lib, ptr

MAX_DEPTH = 10
_PRECISIONS = {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED}


@dataclass(frozen=True)
class FlatOctree:
    depth: np.ndarray
    morton: np.ndarray
    parent: np.ndarray
    child_start: np.ndarray
    child_count: np.ndarray
    point_leaf: np.ndarray

    @property
    def node_count(self) -> int:
        return int(self.depth.size)


@dataclass(frozen=True)
class GridSolution:
    field: np.ndarray
    density: np.ndarray
    screen: np.ndarray
    rhs: np.ndarray
    iso_value: float
    residual_norm: float
    octree: FlatOctree


@dataclass(frozen=True)
class Mesh:
    vertices: np.ndarray
    faces: np.ndarray
    density: np.ndarray
    iso_value: float = 0.0
    residual_norm: float = 0.0

    @property
    def is_empty(self) -> bool:
        return self.vertices.shape[0] == 0


def _points(values) -> np.ndarray:
    result = f64(values)
    if result.ndim != 2 or result.shape[1] != 3:
        raise ValueError("points must have shape (n, 3)")
    return result


def _integer_parameter(name: str, value: int) -> int:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{name} must be an integer")
    return int(value)


def _finite_parameter(name: str, value: float) -> float:
    try:
        result = float(value)
    except (TypeError, ValueError, OverflowError) as error:
        raise TypeError(f"{name} must be a real number") from error
    if not np.isfinite(result):
        raise ValueError(f"{name} must be finite")
    return result


def _precision(value: str) -> int:
    if value not in _PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return _PRECISIONS[value]


def _inner_cycles(value) -> int:
    value = _integer_parameter("inner_cycles", value)
    if value < 0:
        raise ValueError("inner_cycles must be non-negative (0: the library's default)")
    return value


def _valid_oriented_rows(points: np.ndarray, normals: np.ndarray) -> np.ndarray:
    """core.py:73-83: finite rows with a non-zero normal.

    The reference also requires the normal scaled by its largest |component| to have a
    finite, positive squared norm; for a finite normal with a non-zero component that
    squared norm lies in [1, 3], so the test reduces exactly to "all six values finite and
    some normal component non-zero". Written column-wise (no row copies, no reductions
    over the short axis) because it runs on every served request.
    """
    n0, n1, n2 = normals[:, 0], normals[:, 1], normals[:, 2]
    nonzero = (n0 != 0) | (n1 != 0) | (n2 != 0)
    if np.isfinite(points).all() and np.isfinite(normals).all():
        return nonzero
    finite = (np.isfinite(points[:, 0]) & np.isfinite(points[:, 1]) & np.isfinite(points[:, 2])
              & np.isfinite(n0) & np.isfinite(n1) & np.isfinite(n2))
    return finite & nonzero


def _filter(points: np.ndarray, normals: np.ndarray):
    valid = _valid_oriented_rows(points, normals)
    if valid.all():
        return points, normals
    return np.ascontiguousarray(points[valid]), np.ascontiguousarray(normals[valid])


# ---- device jobs ----------------------------------------------------------------------

class Job:
    """One device-resident reconstruction slot (include/mpr_hip.h ``mpr_job``).

    ``host_mesh``: every extraction also copies the mesh into the job's pinned host buffers
    on its stream (``mpr_job_set_host_mesh``), so fetching it is a host copy.
    """

    def __init__(self, depth: int, max_points: int, precision: int = _lib.PRECISION_FP64,
                 host_mesh: bool = True):
        self._lib = lib()
        self.depth = depth
        self.max_points = max_points
        self.precision = precision
        handle = ctypes.c_void_p()
        check(self._lib.mpr_job_create(depth, max_points, precision, ctypes.byref(handle)),
              "mpr_job_create")
        self.handle = handle
        self.count = 0
        self.host_mesh = False
        if host_mesh:
            self.set_host_mesh(True)

    def set_host_mesh(self, enable: bool) -> None:
        check(self._lib.mpr_job_set_host_mesh(self.handle, int(bool(enable))),
              "mpr_job_set_host_mesh")
        self.host_mesh = bool(enable)

    @staticmethod
    def footprint(depth: int, max_points: int, precision: int = _lib.PRECISION_FP64) -> int:
        return int(_lib.load_library().mpr_job_footprint(depth, max_points, precision))

    def close(self) -> None:
        if self.handle:
            self._lib.mpr_job_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, points: np.ndarray, normals: np.ndarray) -> None:
        check(self._lib.mpr_job_set_points(self.handle, ptr(points), ptr(normals), len(points)),
              "mpr_job_set_points")
        self.count = len(points)

    @staticmethod
    def params(stages: int, point_weight=4.0, samples_per_node=1.5, scale=1.1, cycles=4,
               confidence=False, pre_smooth=3, post_smooth=3,
               inner_cycles=0) -> _lib.mpr_params:
        """Run parameters; inner_cycles: FP32 V-cycles per FP64 defect in mixed precision
        (0: the library's default, one)."""
        return _lib.mpr_params(float(point_weight), float(samples_per_node), float(scale),
                               int(cycles), int(bool(confidence)), int(pre_smooth),
                               int(post_smooth), int(stages), int(inner_cycles))

    def run(self, params: _lib.mpr_params) -> None:
        check(self._lib.mpr_job_run(self.handle, ctypes.byref(params)), "mpr_job_run")

    def capture(self, params: _lib.mpr_params) -> None:
        check(self._lib.mpr_job_capture(self.handle, ctypes.byref(params)), "mpr_job_capture")

    def launch(self) -> None:
        check(self._lib.mpr_job_launch(self.handle), "mpr_job_launch")

    def done(self) -> bool:
        status = self._lib.mpr_job_query(self.handle)
        if status < 0:
            raise _lib.DeviceError("mpr_job_query: " + _lib.last_error())
        return status == 1

    def wait(self) -> _lib.mpr_result:
        result = _lib.mpr_result()
        check(self._lib.mpr_job_wait(self.handle, ctypes.byref(result)), "mpr_job_wait")
        return result

    def fetch_mesh(self, result: _lib.mpr_result):
        nv, nf = int(result.vertex_count), int(result.face_count)
        vertices = np.empty((nv, 3), np.float64)
        density = np.empty(nv, np.float64)
        faces = np.empty((nf, 3), np.int64)
        check(self._lib.mpr_job_fetch_mesh(self.handle, ptr(vertices), ptr(density), ptr(faces)),
              "mpr_job_fetch_mesh")
        return vertices, faces, density

    def fetch_grids(self):
        """[field, density, screen, rhs] of the last run, which needs STAGE_GRIDS."""
        n = (1 << self.depth) + 1
        grids = [np.empty((n, n, n), np.float64) for _ in range(4)]
        check(self._lib.mpr_job_fetch_grids(self.handle, *(ptr(g) for g in grids)),
              "mpr_job_fetch_grids")
        return grids

    def octree(self, depth: int) -> FlatOctree:
        count = ctypes.c_int64()
        check(self._lib.mpr_job_build_octree(self.handle, depth, ctypes.byref(count)),
              "mpr_job_build_octree")
        nodes = int(count.value)
        arrays = [np.empty(nodes, np.int64) for _ in range(5)]
        leaf = np.empty(self.count, np.int64)
        check(self._lib.mpr_job_fetch_octree(self.handle, *(ptr(a) for a in arrays), ptr(leaf)),
              "mpr_job_fetch_octree")
        return FlatOctree(*arrays, leaf)


class _JobCache:
    """Idle device jobs by (depth, precision), each lent to one caller at a time: repeated
    calls do not re-allocate HBM, and concurrent calls (from several threads) never share a
    job, as the reference's stateless calls never share buffers."""

    def __init__(self):
        self._idle: dict[tuple[int, int], Job] = {}
        self._lock = threading.Lock()

    @contextmanager
    def lend(self, depth: int, count: int, precision: int):
        key = (depth, precision)
        with self._lock:
            job = self._idle.pop(key, None)
        if job is not None and job.max_points < count:
            job.close()
            job = None
        if job is None:
            job = Job(depth, max(count, 1), precision)
        try:
            yield job
        except BaseException:
            job.close()  # its device state is unknown after a failure
            raise
        with self._lock:
            other = self._idle.get(key)
            if other is None or other.max_points < job.max_points:
                self._idle[key] = job
                job = other
        if job is not None:  # keep the larger of two returned jobs
            job.close()

    def clear(self) -> None:
        with self._lock:
            for job in self._idle.values():
                job.close()
            self._idle.clear()


_jobs = _JobCache()


def release_device_memory() -> None:
    """Free the cached device jobs."""
    _jobs.clear()


# ---- public API ---------------------------------------------------------------------------

def build_octree(points, depth: int = 6) -> FlatOctree:
    """Build a compact breadth-first Morton octree over unit-cube points (core.py:86-127)."""
    points = _points(points)
    depth = _integer_parameter("depth", depth)
    if not 0 <= depth <= 19:
        raise ValueError("depth must be between 0 and 19")
    if len(points) and (not np.isfinite(points).all() or np.any(points < 0) or np.any(points > 1)):
        raise ValueError("octree points must be finite and inside [0, 1]^3")
    if not len(points):
        one = np.array([0], dtype=np.int64)
        zero = np.array([0], dtype=np.int64)
        return FlatOctree(zero, one, np.array([-1], np.int64), np.array([-1], np.int64), zero,
                          np.empty(0, np.int64))
    with _jobs.lend(1, len(points), _lib.PRECISION_FP64) as job:
        job.set_points(points, points)
        return job.octree(depth)


def solve_screened_poisson(
    points,
    normals,
    *,
    depth: int = 6,
    point_weight: float = 4.0,
    confidence: bool = False,
    cycles: int = 4,
    samples_per_node: float = 1.5,
    precision: str = "fp64",
    inner_cycles: int = 0,
) -> GridSolution:
    """Build the octree, splat samples, and solve the screened Poisson system (core.py:196-306).
    precision and inner_cycles (mixed precision: FP32 V-cycles per FP64 defect, 0 for the
    library's default) have no counterpart in the reference."""
    points = _points(points)
    normals = _points(normals)
    if points.shape != normals.shape:
        raise ValueError("points and normals must have the same shape")
    depth = _integer_parameter("depth", depth)
    cycles = _integer_parameter("cycles", cycles)
    point_weight = _finite_parameter("point_weight", point_weight)
    samples_per_node = _finite_parameter("samples_per_node", samples_per_node)
    mode = _precision(precision)
    inner_cycles = _inner_cycles(inner_cycles)
    if not 2 <= depth <= MAX_DEPTH:
        raise ValueError(f"depth must be between 2 and {MAX_DEPTH}")
    if cycles < 1:
        raise ValueError("cycles must be positive")
    if point_weight < 0 or samples_per_node <= 0:
        raise ValueError("point_weight must be non-negative and samples_per_node positive")

    points, normals = _filter(points, normals)
    if len(points) and (np.any(points < 0) or np.any(points > 1)):
        raise ValueError("solver points must be inside [0, 1]^3")

    with _jobs.lend(depth, len(points), mode) as job:
        job.set_points(points, normals)
        job.run(Job.params(_lib.STAGE_SOLVE | _lib.STAGE_GRIDS, point_weight, samples_per_node,
                           cycles=cycles, confidence=confidence, inner_cycles=inner_cycles))
        result = job.wait()
        if result.accepted != len(points):
            raise RuntimeError(
                f"GPU splat accepted {result.accepted} of {len(points)} validated samples")
        field, density, screen, rhs = job.fetch_grids()
        tree = job.octree(depth) if len(points) else build_octree(points, depth)
    return GridSolution(field, density, screen, rhs, float(result.iso_value),
                        float(result.residual_norm), tree)


def extract_isosurface(field, *, iso_value: float = 0.0, density=None) -> Mesh:
    """Extract a triangle mesh from a cubic scalar grid on the dual grid (core.py:309-374)."""
    field = f64(field)
    if field.ndim != 3 or len(set(field.shape)) != 1 or field.shape[0] < 2:
        raise ValueError("field must be a cubic array with side length >= 2")
    iso_value = _finite_parameter("iso_value", iso_value)
    if not np.isfinite(field).all():
        raise ValueError("field and iso_value must be finite")
    density = np.zeros_like(field) if density is None else f64(density)
    if density.shape != field.shape:
        raise ValueError("density must have the same shape as field")
    if not np.isfinite(density).all():
        raise ValueError("density must be finite")
    n = field.shape[0]
    handle = ctypes.c_void_p()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    check(lib().mpr_extract(addr(field), addr(density), n, iso_value, ctypes.byref(handle),
                            ctypes.byref(nv), ctypes.byref(nf)), "mpr_extract")
    try:
        count, face_count = int(nv.value), int(nf.value)
        if count < 0 or count > (n - 1) ** 3:
            raise RuntimeError(f"invalid dual vertex count returned by the GPU: {count}")
        if face_count < 0 or face_count > 6 * (n - 1) ** 3:
            raise RuntimeError(f"invalid dual face count returned by the GPU: {face_count}")
        if count == 0:
            return Mesh(np.empty((0, 3), np.float64), np.empty((0, 3), np.int64), np.empty(0),
                        iso_value)
        vertices = np.empty((count, 3), np.float64)
        vertex_density = np.empty(count, np.float64)
        faces = np.empty((face_count, 3), np.int64)
        check(lib().mpr_mesh_fetch(handle, ptr(vertices), ptr(vertex_density), ptr(faces)),
              "mpr_mesh_fetch")
    finally:
        lib().mpr_mesh_destroy(handle)
    return Mesh(vertices, faces, vertex_density, iso_value)


def reconstruct(
    points,
    normals,
    *,
    depth: int = 6,
    point_weight: float = 4.0,
    confidence: bool = False,
    cycles: int = 4,
    scale: float = 1.1,
    samples_per_node: float = 1.5,
    precision: str = "fp64",
    inner_cycles: int = 0,
) -> Mesh:
    """Reconstruct a watertight surface from oriented points (core.py:388-426). precision and
    inner_cycles as solve_screened_poisson's."""
    req = prepare_reconstruct(points, normals, depth=depth, point_weight=point_weight,
                              confidence=confidence, cycles=cycles, scale=scale,
                              samples_per_node=samples_per_node, precision=precision,
                              inner_cycles=inner_cycles)
    if req is None:
        return _empty_mesh()
    with _jobs.lend(req.depth, len(req.points), req.precision) as job:
        job.set_points(req.points, req.normals)
        job.run(req.params)
        return mesh_from_job(job, job.wait(), len(req.points))


@dataclass(frozen=True)
class ReconstructRequest:
    """A validated reconstruct() call: its rows (filtered on the device), depth, precision
    and device params."""
    points: np.ndarray
    normals: np.ndarray
    depth: int
    precision: int
    params: _lib.mpr_params


def prepare_reconstruct(points, normals, *, depth=6, point_weight=4.0, confidence=False,
                        cycles=4, scale=1.1, samples_per_node=1.5, precision="fp64",
                        inner_cycles=0):
    """Validation of reconstruct() (core.py:400-411 and the solver checks of
    core.py:211-220), in the reference's order. The row filter of core.py:407-409 runs on the
    device (MPR_STAGE_FILTER), fused with the normalisation. Returns None when there is no
    row at all (or, as in the reference, when no row survives the filter and the solver
    parameters were never looked at)."""
    points = _points(points)
    normals = _points(normals)
    if points.shape != normals.shape:
        raise ValueError("points and normals must have the same shape")
    scale = _finite_parameter("scale", scale)
    if scale <= 1.0:
        raise ValueError("scale must be greater than 1")
    if not len(points):
        return None
    try:
        depth = _integer_parameter("depth", depth)
        cycles = _integer_parameter("cycles", cycles)
        point_weight = _finite_parameter("point_weight", point_weight)
        samples_per_node = _finite_parameter("samples_per_node", samples_per_node)
        mode = _precision(precision)
        inner_cycles = _inner_cycles(inner_cycles)
        if not 2 <= depth <= MAX_DEPTH:
            raise ValueError(f"depth must be between 2 and {MAX_DEPTH}")
        if cycles < 1:
            raise ValueError("cycles must be positive")
        if point_weight < 0 or samples_per_node <= 0:
            raise ValueError("point_weight must be non-negative and samples_per_node positive")
    except (TypeError, ValueError):
        # the reference validates these only once a row has survived its filter
        # (core.py:407-421): with none left it returns an empty mesh instead of raising
        if not _valid_oriented_rows(points, normals).any():
            return None
        raise
    stages = (_lib.STAGE_FILTER | _lib.STAGE_NORMALIZE | _lib.STAGE_SOLVE | _lib.STAGE_EXTRACT
              | _lib.STAGE_WORLD)
    params = Job.params(stages, point_weight, samples_per_node, scale, cycles, confidence,
                        inner_cycles=inner_cycles)
    return ReconstructRequest(points, normals, depth, mode, params)


def _empty_mesh() -> Mesh:
    return Mesh(np.empty((0, 3), np.float64), np.empty((0, 3), np.int64), np.empty(0))


def mesh_from_job(job: Job, result: _lib.mpr_result, count: int | None = None) -> Mesh:
    """The world-space Mesh of a finished reconstruct run (core.py:410-411, 422-426). The
    splat must have accepted every row the device filter kept (`count`, when given, is the
    number of rows uploaded)."""
    kept = int(result.valid_rows)
    if count is not None and kept > count:
        raise RuntimeError(f"GPU filter kept {kept} of {count} rows")
    if result.accepted != kept:
        raise RuntimeError(f"GPU splat accepted {result.accepted} of {kept} validated samples")
    if kept == 0:  # core.py:410-411: no valid row
        return _empty_mesh()
    vertices, faces, density = job.fetch_mesh(result)
    if not len(vertices):
        return Mesh(np.empty((0, 3), np.float64), np.empty((0, 3), np.int64), np.empty(0),
                    float(result.iso_value), float(result.residual_norm))
    return Mesh(vertices, faces, density, float(result.iso_value), float(result.residual_norm))


def oriented_points_from_mesh(vertices, faces) -> tuple[np.ndarray, np.ndarray]:
    """Referenced vertices and area-weighted normals of a triangle mesh (core.py:429-450).

    Computed on the GPU (mpr_oriented_points_from_mesh) with the arithmetic of the
    reference's per-face loop: unnormalised cross products added to each corner in face
    order, then normalised, so the result is the same bit for bit.
    """
    vertices = _points(vertices)
    faces = i64(faces)
    if faces.ndim != 2 or faces.shape[1] != 3:
        raise ValueError("faces must have shape (m, 3)")
    if len(faces) and (faces.min() < 0 or faces.max() >= len(vertices)):
        raise ValueError("face index out of range")
    points = np.empty_like(vertices)
    normals = np.empty_like(vertices)
    count = ctypes.c_int64()
    check(lib().mpr_oriented_points_from_mesh(ptr(vertices), len(vertices), ptr(faces),
                                              len(faces), ptr(points), ptr(normals),
                                              ctypes.byref(count)),
          "mpr_oriented_points_from_mesh")
    k = int(count.value)
    return np.ascontiguousarray(points[:k]), np.ascontiguousarray(normals[:k])


def reconstruct_mesh(vertices, faces, **kwargs) -> Mesh:
    points, normals = oriented_points_from_mesh(vertices, faces)
    return reconstruct(points, normals, **kwargs)
