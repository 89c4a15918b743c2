"""Batched reconstruction: many independent point clouds in one device pass.

The reference reconstructs one cloud per call. Here a batch of jobs (one per cloud, of any
depths) runs its V-cycles together: every V-cycle step is one kernel launch over all jobs
that have a grid of that size, and the whole batched run can be recorded once as a HIP
graph and replayed with a single launch. Each cloud's arithmetic is exactly what its solo
reconstruct() performs (every stage is deterministic), so results are bit-identical to solo
runs.
"""

from __future__ import annotations

import ctypes
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from ._lib import check, lib
from .core import Job, Mesh, mesh_from_job, prepare_reconstruct


_pool = None
_pool_lock = threading.Lock()


def host_pool() -> ThreadPoolExecutor:
    """Threads for the host side of a batch: uploads and mesh copies release the GIL in the
    library (hipMemcpy staging, memcpy out of pinned memory and the page faults of fresh
    NumPy arrays), so several run at once."""
    global _pool
    with _pool_lock:
        if _pool is None:
            _pool = ThreadPoolExecutor(max_workers=8, thread_name_prefix="mpr-host")
        return _pool


def upload_all(jobs, clouds) -> None:
    """set_points of every (job, (points, normals)) pair, on the host pool."""
    list(host_pool().map(lambda jc: jc[0].set_points(*jc[1]), zip(jobs, clouds)))


def meshes_from_jobs(jobs, results, counts=None) -> list[Mesh]:
    """mesh_from_job of every job of a finished batch, on the host pool."""
    counts = [None] * len(jobs) if counts is None else counts
    return list(host_pool().map(lambda a: mesh_from_job(*a), zip(jobs, results, counts)))


class Batch:
    """A fixed set of jobs whose V-cycles run as one schedule (include/mpr_hip.h mpr_batch)."""

    def __init__(self, jobs: list[Job]):
        if not jobs:
            raise ValueError("a batch needs at least one job")
        self._lib = lib()
        self.jobs = list(jobs)
        arr = (ctypes.c_void_p * len(jobs))(*(j.handle for j in jobs))
        handle = ctypes.c_void_p()
        check(self._lib.mpr_batch_create(arr, len(jobs), ctypes.byref(handle)), "mpr_batch_create")
        self.handle = handle

    def close(self) -> None:
        if self.handle:
            self._lib.mpr_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _params(self, params) -> ctypes.Array:
        if isinstance(params, _lib.mpr_params):
            params = [params] * len(self.jobs)
        if len(params) != len(self.jobs):
            raise ValueError("one parameter set per job is required")
        return (_lib.mpr_params * len(params))(*params)

    def run(self, params) -> None:
        check(self._lib.mpr_batch_run(self.handle, self._params(params)), "mpr_batch_run")

    def capture(self, params) -> None:
        check(self._lib.mpr_batch_capture(self.handle, self._params(params)), "mpr_batch_capture")

    def launch(self) -> None:
        check(self._lib.mpr_batch_launch(self.handle), "mpr_batch_launch")

    def done(self) -> bool:
        status = self._lib.mpr_batch_query(self.handle)
        if status < 0:
            raise _lib.DeviceError("mpr_batch_query: " + _lib.last_error())
        return status == 1

    def wait(self) -> list[_lib.mpr_result]:
        results = (_lib.mpr_result * len(self.jobs))()
        check(self._lib.mpr_batch_wait(self.handle, results), "mpr_batch_wait")
        return list(results)


def reconstruct_batch(clouds, *, depths, point_weight: float = 4.0, confidence=False,
                      cycles: int = 4, scale: float = 1.1, samples_per_node: float = 1.5,
                      precision: str = "mixed", graph: bool = False,
                      inner_cycles: int = 0) -> list[Mesh]:
    """Reconstruct each ``(points, normals)`` cloud at its depth with one batched device pass.

    ``depths`` and ``confidence`` may be scalars or per-cloud sequences. Every cloud gets the
    validation of reconstruct(); a cloud with no valid rows yields an empty Mesh. With
    ``graph`` the batched run is recorded as one HIP graph and replayed as a single launch.
    """
    clouds = list(clouds)
    count = len(clouds)
    depths = [depths] * count if np.isscalar(depths) else list(depths)
    conf = [confidence] * count if np.isscalar(confidence) else list(confidence)
    if len(depths) != count or len(conf) != count:
        raise ValueError("depths and confidence must match the number of clouds")
    reqs = [prepare_reconstruct(p, n, depth=d, point_weight=point_weight, confidence=c,
                                cycles=cycles, scale=scale, samples_per_node=samples_per_node,
                                precision=precision, inner_cycles=inner_cycles)
            for (p, n), d, c in zip(clouds, depths, conf)]
    live = [i for i, r in enumerate(reqs) if r is not None]
    meshes: list[Mesh] = [Mesh(np.empty((0, 3)), np.empty((0, 3), np.int64), np.empty(0))
                          for _ in range(count)]
    if not live:
        return meshes
    jobs = [Job(reqs[i].depth, len(reqs[i].points), reqs[i].precision) for i in live]
    try:
        upload_all(jobs, [(reqs[i].points, reqs[i].normals) for i in live])
        batch = Batch(jobs)
        try:
            params = [reqs[i].params for i in live]
            if graph:
                batch.capture(params)  # records only
                batch.launch()
            else:
                batch.run(params)
            results = batch.wait()
        finally:
            batch.close()
        for i, m in zip(live, meshes_from_jobs(jobs, results,
                                               [len(reqs[i].points) for i in live])):
            meshes[i] = m
    finally:
        for j in jobs:
            j.close()
    return meshes
