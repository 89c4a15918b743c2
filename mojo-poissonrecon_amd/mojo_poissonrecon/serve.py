"""Continuous-batching request loop for one GPU.

Requests (oriented point clouds with a depth) are admitted while device memory and slots
allow: an admitted request takes an idle device job of its depth from a pool (or a new
one, if its HBM footprint fits the budget), uploads its points on that job's stream and
enqueues the whole reconstruction without waiting. Requests of one depth admitted in the
same tick run as one batch (one launch per V-cycle step for all of them). Work of
different requests overlaps on the device; completed meshes are read back as soon as their
stream finishes, which frees the slot and memory for the next arrivals. Latency is
measured from a request's arrival to its mesh being on the host.

The FP64 single-request reconstruct() stays the reference path; each served result is
bit-identical to that request's solo run (every stage is deterministic and an instance's
arithmetic does not depend on its batch).
"""

from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .batch import Batch
from .core import Job, Mesh, mesh_from_job, prepare_reconstruct


@dataclass
class Request:
    id: int
    points: np.ndarray
    normals: np.ndarray
    depth: int
    params: _lib.mpr_params
    precision: int
    arrival: float
    admitted: float = 0.0
    completed: float = 0.0
    mesh: Mesh | None = None

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


@dataclass
class _Flight:
    requests: list
    jobs: list
    batch: Batch | None = None


@dataclass
class ServerStats:
    completed: int = 0
    batches: int = 0
    peak_inflight: int = 0
    peak_bytes: int = 0
    jobs_created: int = 0
    latencies: list = field(default_factory=list)
    # host seconds spent in each part of the loop (validation, upload + launch, readback)
    host_prepare: float = 0.0
    host_launch: float = 0.0
    host_collect: float = 0.0


def job_capacity(points: int) -> int:
    """Point capacity of a pooled job: the next power of two, at least 16384, so that a
    pool job serves later clouds of similar size instead of only the one it was made for."""
    return max(16384, 1 << max(0, int(points) - 1).bit_length())


class ReconstructionServer:
    """Admits reconstruction requests by HBM footprint and overlaps them on the device."""

    def __init__(self, *, hbm_budget: int | None = None, max_inflight: int = 12,
                 max_batch: int = 8, precision: str = "mixed"):
        self.precision = precision
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        if hbm_budget is None:
            total = _lib.ctypes.c_int64()
            _lib.check(_lib.lib().mpr_device_info(None, 0, _lib.ctypes.byref(total), None),
                       "mpr_device_info")
            hbm_budget = int(0.7 * total.value)
        self.hbm_budget = hbm_budget
        self._pool: list[Job] = []        # idle jobs
        self._busy: set[int] = set()      # ids of jobs in flight
        self._bytes = 0
        self._waiting: list[Request] = []
        self._flights: list[_Flight] = []
        self._ids = itertools.count()
        self.stats = ServerStats()

    # ---- public ---------------------------------------------------------------------
    def submit(self, points, normals, *, depth: int, arrival: float | None = None,
               **kwargs) -> Request | None:
        kwargs.setdefault("precision", self.precision)
        t0 = time.perf_counter()
        req = prepare_reconstruct(points, normals, depth=depth, **kwargs)
        self.stats.host_prepare += time.perf_counter() - t0
        now = t0 if arrival is None else arrival
        if req is None:
            return None
        r = Request(next(self._ids), req.points, req.normals, req.depth, req.params,
                    req.precision, now)
        self._waiting.append(r)
        return r

    def step(self) -> list[Request]:
        """Admit what fits, launch it, and collect finished requests (non-blocking)."""
        t0 = time.perf_counter()
        done = self._collect()
        t1 = time.perf_counter()
        self._admit()
        self.stats.host_collect += t1 - t0
        self.stats.host_launch += time.perf_counter() - t1
        return done

    def drain(self) -> list[Request]:
        out = []
        while self._waiting or self._flights:
            out += self.step()
            if self._flights and not out:
                time.sleep(0)
        return out

    def close(self) -> None:
        for f in self._flights:
            if f.batch is not None:
                f.batch.wait()
            else:
                for j in f.jobs:
                    j.wait()
        for f in self._flights:
            if f.batch is not None:
                f.batch.close()
        self._flights.clear()
        for j in self._pool:
            j.close()
        self._pool.clear()

    @property
    def inflight(self) -> int:
        return sum(len(f.requests) for f in self._flights)

    # ---- internals ------------------------------------------------------------------
    def _take_job(self, r: Request) -> Job | None:
        best = None
        for j in self._pool:
            if (id(j) not in self._busy and j.depth == r.depth and j.precision == r.precision
                    and j.max_points >= len(r.points)):
                if best is None or j.max_points < best.max_points:
                    best = j
        if best is not None:
            return best
        capacity = job_capacity(len(r.points))
        need = Job.footprint(r.depth, capacity, r.precision)
        # evict idle jobs until the new one fits the budget
        idle = [j for j in self._pool if id(j) not in self._busy]
        idle.sort(key=lambda j: Job.footprint(j.depth, j.max_points, j.precision))
        while self._bytes + need > self.hbm_budget and idle:
            victim = idle.pop()
            self._bytes -= Job.footprint(victim.depth, victim.max_points, victim.precision)
            self._pool.remove(victim)
            victim.close()
        if self._bytes + need > self.hbm_budget:
            return None
        job = Job(r.depth, capacity, r.precision)
        self.stats.jobs_created += 1
        self._pool.append(job)
        self._bytes += need
        self.stats.peak_bytes = max(self.stats.peak_bytes, self._bytes)
        return job

    def _admit(self) -> None:
        # one batch per depth: a batch completes as a whole, so shallow requests are not
        # held back by a deeper one admitted in the same tick
        groups: dict[int, list[tuple[Request, Job]]] = {}
        pending = 0
        while self._waiting and self.inflight + pending < self.max_inflight:
            r = self._waiting[0]
            job = self._take_job(r)
            if job is None:
                break  # memory is full: wait for a completion (FIFO admission)
            self._waiting.pop(0)
            self._busy.add(id(job))
            job.set_points(r.points, r.normals)
            r.admitted = time.perf_counter()
            group = groups.setdefault(r.depth, [])
            group.append((r, job))
            pending += 1
            if len(group) == self.max_batch:
                self._launch(group)
                pending -= len(group)
                groups[r.depth] = []
        for group in groups.values():
            if group:
                self._launch(group)
        self.stats.peak_inflight = max(self.stats.peak_inflight, self.inflight)

    def _launch(self, group) -> None:
        reqs = [r for r, _ in group]
        jobs = [j for _, j in group]
        uniform = all(r.precision == reqs[0].precision for r in reqs) and all(
            (r.params.cycles, r.params.pre_smooth, r.params.post_smooth, r.params.point_weight)
            == (reqs[0].params.cycles, reqs[0].params.pre_smooth, reqs[0].params.post_smooth,
                reqs[0].params.point_weight) for r in reqs)
        if len(group) > 1 and uniform:
            batch = Batch(jobs)
            batch.run([r.params for r in reqs])
            self._flights.append(_Flight(reqs, jobs, batch))
            self.stats.batches += 1
        else:
            for r, j in group:
                j.run(r.params)
                self._flights.append(_Flight([r], [j]))

    def _collect(self) -> list[Request]:
        done = []
        still = []
        for f in self._flights:
            ready = f.batch.done() if f.batch is not None else f.jobs[0].done()
            if not ready:
                still.append(f)
                continue
            results = f.batch.wait() if f.batch is not None else [f.jobs[0].wait()]
            for r, j, res in zip(f.requests, f.jobs, results):
                r.mesh = mesh_from_job(j, res, len(r.points))
                r.completed = time.perf_counter()
                self._busy.discard(id(j))
                self.stats.completed += 1
                self.stats.latencies.append(r.latency)
                done.append(r)
            if f.batch is not None:
                f.batch.close()
        self._flights = still
        return done
