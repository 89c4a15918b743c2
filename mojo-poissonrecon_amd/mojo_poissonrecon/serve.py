"""Continuous-batching request loop for one GPU.

Requests (oriented point clouds with a depth) are admitted while device memory and slots
allow: an admitted request takes an idle device job of its depth from a pool (or a new
one, if its HBM footprint fits the budget), uploads its raw rows on that job's stream (the
row filter and normalisation run on the device) and enqueues the whole reconstruction
without waiting. Requests of one depth admitted in the same tick run as one batch (one
launch per V-cycle step for all of them); batches over the same pooled jobs are reused.
Work of different requests overlaps on the device; each extraction copies its mesh into
its job's pinned host buffers on the job's stream, so a finished request's mesh is read
back with a host copy. Uploads, mesh copies, the creation of new pool jobs and the release
of evicted ones (device allocations and frees, which can wait behind the work in flight)
run on a thread pool and the loop never waits for them: the requests admitted in one tick
are launched once all their jobs are made and uploaded, and a finished flight's requests
complete once their copies are. Latency is measured from a request's arrival to its mesh
being on the host.

Failures are isolated per flight (one batch, or one solo request): its jobs are discarded,
the rest of the server keeps running, and each of its requests is run again alone as the
plain FP64 reconstruct() (the reference path) or, if that fails too, completed with its
error. Each served result is bit-identical to that request's solo run (every stage is
deterministic and an instance's arithmetic does not depend on its batch).
"""

from __future__ import annotations

import functools
import itertools
import time
from collections import OrderedDict
from concurrent.futures import Future, ThreadPoolExecutor
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
    kwargs: dict = field(default_factory=dict)
    admitted: float = 0.0
    completed: float = 0.0
    mesh: Mesh | None = None
    error: str | None = None   # set when the request failed, fallback included
    fallback: bool = False     # served by the FP64 single-request path after a failure
    # with the server's verify_every: resolves to {"hausdorff", "chamfer"}, the exact
    # distances between this mesh and the verifier's (plain FP64) mesh over the bbox diagonal
    fp64_error: Future | None = None

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


@dataclass
class _Flight:
    requests: list
    jobs: list
    batch: Batch | None = None
    copies: list | None = None  # the mesh copies' futures, once the device work is done


@dataclass
class ServerStats:
    completed: int = 0
    failed: int = 0
    fallbacks: int = 0
    batches: int = 0
    batches_reused: int = 0
    peak_inflight: int = 0
    peak_bytes: int = 0
    jobs_created: int = 0
    verified: int = 0          # results sent to the verifier (verify_every)
    latencies: list = field(default_factory=list)
    # host seconds spent in each part of the loop (validation, upload + launch, readback)
    host_prepare: float = 0.0
    host_launch: float = 0.0
    host_collect: float = 0.0


def job_capacity(points: int, depth: int | None = None,
                 precision: int = _lib.PRECISION_MIXED) -> int:
    """Point capacity of a pooled job: the next power of two, at least 16384, so that a pool
    job serves later clouds of similar size instead of only the one it was made for. With
    `depth`, raised further while the job's HBM footprint stays within 3 % of the depth's
    smallest job: on deep grids the point buffers are a small part of it (a depth-9 job for
    512k points is 1.6 % larger than one for 16k), and one capacity class lets every pooled
    job of that depth serve every such request instead of a new job being made."""
    c = max(16384, 1 << max(0, int(points) - 1).bit_length())
    return c if depth is None else _grown_capacity(c, int(depth), int(precision))


@functools.lru_cache(maxsize=None)
def _grown_capacity(c: int, depth: int, precision: int) -> int:
    limit = 1.03 * Job.footprint(depth, 16384, precision)
    while c < (1 << 24) and Job.footprint(depth, 2 * c, precision) <= limit:
        c *= 2
    return c


def _fp64_fallback(r: Request) -> Mesh:
    from .core import reconstruct

    kw = dict(r.kwargs)
    kw["precision"] = "fp64"
    return reconstruct(r.points, r.normals, depth=r.depth, **kw)


class ReconstructionServer:
    """Admits reconstruction requests by HBM footprint and overlaps them on the device.

    ``job_factory(depth, capacity, precision)``, ``batch_factory(jobs)`` and
    ``fallback(request) -> Mesh`` default to the device Job, Batch and the FP64 solo
    reconstruct(); they can be replaced (tests use stand-ins to force failures).

    ``verify_every=k`` (0 = off) sends every k-th completed request (by id) to the verifier,
    ``verifier(request) -> Mesh`` (default: the same plain FP64 reconstruct), on the host
    thread pool, and sets ``request.fp64_error`` to a future of the exact Hausdorff and
    Chamfer distances between the two meshes over the bbox diagonal of the request's
    finite rows (``distance(a, b)``, default ``metrics.mesh_distance``). The request
    completes without waiting for it; the verifier's solves share the device with the
    stream.
    """

    def __init__(self, *, hbm_budget: int | None = None, max_inflight: int = 12,
                 max_batch: int = 8, precision: str = "mixed", host_threads: int = 8,
                 batch_cache: int = 8, job_factory=None, batch_factory=None,
                 fallback=_fp64_fallback, linger: float = 0.004, min_active: int = 2,
                 verify_every: int = 0, verifier=None, distance=None):
        if not isinstance(verify_every, int) or verify_every < 0:
            raise ValueError("verify_every must be a non-negative int")
        self.precision = precision
        self.verify_every = verify_every
        self._verifier = verifier or _fp64_fallback
        self._distance = distance
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        # uploaded requests of one depth wait to be launched together: until max_batch of
        # them are ready, the device has fewer than min_active flights running, or the
        # oldest has waited `linger` seconds since its admission
        self.linger = linger
        self.min_active = min_active
        if hbm_budget is None:
            total = _lib.ctypes.c_int64()
            _lib.check(_lib.lib().mpr_device_info(None, 0, _lib.ctypes.byref(total), None),
                       "mpr_device_info")
            hbm_budget = int(0.7 * total.value)
        self.hbm_budget = hbm_budget
        self._job_factory = job_factory or (lambda d, c, p: Job(d, c, p))
        self._batch_factory = batch_factory or Batch
        self._fallback = fallback
        self._pool: list = []             # jobs (idle or busy)
        self._busy: set[int] = set()      # ids of jobs in flight
        self._bytes = 0
        self._waiting: list[Request] = []
        self._flights: list[_Flight] = []
        # (request, upload future) admitted in one tick; the future makes the request's job
        # first when no idle one fitted, and returns (job, error)
        self._cohorts: list[list] = []
        self._staged: dict[int, list] = {}  # depth -> (request, job) uploaded, not launched
        self._ready: list[Request] = []   # failed flights' requests, returned by the next step
        self._batches: OrderedDict = OrderedDict()  # job ids -> Batch, least recent first
        self._batch_cache = batch_cache
        self._threads = ThreadPoolExecutor(max_workers=max(1, host_threads))
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
                    req.precision, now, kwargs={k: v for k, v in kwargs.items()
                                                if k != "precision"})
        self._waiting.append(r)
        return r

    def step(self) -> list[Request]:
        """Collect finished requests, then admit and launch what fits (non-blocking)."""
        t0 = time.perf_counter()
        done = self._collect()
        t1 = time.perf_counter()
        self._admit()
        self.stats.host_collect += t1 - t0
        self.stats.host_launch += time.perf_counter() - t1
        return done

    def drain(self) -> list[Request]:
        out = []
        while self._waiting or self._flights or self._ready or self._cohorts or self._staged:
            out += self.step()
            if (self._flights or self._cohorts or self._staged) and not out:
                time.sleep(0)
        return out

    def close(self) -> None:
        for cohort in self._cohorts:
            for _, fut in cohort:
                job, _ = fut.result()  # an upload still copying from the caller's rows
                if job is not None and job not in self._pool:
                    self._pool.append(job)
        self._cohorts.clear()
        self._staged.clear()
        for f in self._flights:
            for c in f.copies or ():
                try:
                    c.result()
                except Exception:
                    pass
            try:
                if f.batch is not None:
                    f.batch.wait()
                else:
                    for j in f.jobs:
                        j.wait()
            except Exception:  # the jobs are closed below either way
                pass
        self._flights.clear()
        for b in self._batches.values():
            b.close()
        self._batches.clear()
        for j in self._pool:
            j.close()
        self._pool.clear()
        self._busy.clear()
        self._bytes = 0
        self._threads.shutdown(wait=True)

    @property
    def inflight(self) -> int:
        return (sum(len(f.requests) for f in self._flights) + sum(len(c) for c in self._cohorts)
                + sum(len(g) for g in self._staged.values()))

    # ---- internals ------------------------------------------------------------------
    def _footprint(self, j) -> int:
        return Job.footprint(j.depth, j.max_points, j.precision)

    def _detach_job(self, job) -> list:
        """Take a job and its batches out of the server; returns them, to be closed."""
        gone = [self._batches.pop(k) for k in [k for k in self._batches if id(job) in k]]
        if job in self._pool:
            self._pool.remove(job)
            self._bytes -= self._footprint(job)
        self._busy.discard(id(job))
        return gone + [job]

    @staticmethod
    def _close_all(objs) -> None:
        for o in objs:
            try:
                o.close()
            except Exception:
                pass

    def _discard_job(self, job) -> None:
        """Drop a job in an unknown state after a failure, and its batches."""
        self._close_all(self._detach_job(job))

    def _take_job(self, r: Request):
        """An idle pooled job for `r`, or (capacity, footprint, evicted) for a new one that
        fits the budget once the idle jobs in `evicted` are gone, or None."""
        best = None
        for j in self._pool:
            if (id(j) not in self._busy and j.depth == r.depth and j.precision == r.precision
                    and j.max_points >= len(r.points)):
                if best is None or j.max_points < best.max_points:
                    best = j
        if best is not None:
            return best
        capacity = job_capacity(len(r.points), r.depth, r.precision)
        need = Job.footprint(r.depth, capacity, r.precision)
        # evict idle jobs (largest first) until the new one fits the budget
        idle = [j for j in self._pool if id(j) not in self._busy]
        idle.sort(key=self._footprint)
        evicted = []
        while self._bytes + need > self.hbm_budget and idle:
            evicted += self._detach_job(idle.pop())
        if self._bytes + need > self.hbm_budget:
            self._threads.submit(self._close_all, evicted)
            return None
        self.stats.jobs_created += 1
        self._bytes += need  # held from now; the job is made on the thread pool
        self.stats.peak_bytes = max(self.stats.peak_bytes, self._bytes)
        return capacity, need, evicted

    def _admit(self) -> None:
        # one batch per depth: a batch completes as a whole, so shallow requests are not
        # held back by a deeper one admitted in the same tick
        admitted: list = []
        while self._waiting and self.inflight + len(admitted) < self.max_inflight:
            r = self._waiting[0]
            job = self._take_job(r)
            if job is None:
                break  # memory is full: wait for a completion (FIFO admission)
            self._waiting.pop(0)
            if not isinstance(job, tuple):
                self._busy.add(id(job))
            admitted.append((r, job))
        if admitted:
            # job creation and uploads on the thread pool (an upload is a host copy that
            # releases the GIL)
            now = time.perf_counter()
            cohort = []
            for r, job in admitted:
                r.admitted = now
                cohort.append((r, self._threads.submit(self._upload, r, job)))
            self._cohorts.append(cohort)
        # stage the ticks whose uploads are all done (a tick's requests stay together), in
        # admission order
        while self._cohorts and all(fut.done() for _, fut in self._cohorts[0]):
            for r, fut in self._cohorts.pop(0):
                job, err = fut.result()
                if job is not None and job not in self._pool:
                    self._pool.append(job)
                    self._busy.add(id(job))
                if err is not None:
                    if job is None:  # its creation failed: give back the bytes it held
                        self._bytes -= Job.footprint(
                            r.depth, job_capacity(len(r.points), r.depth, r.precision),
                            r.precision)
                    self._fail(_Flight([r], [job] if job is not None else []), err)
                else:
                    self._staged.setdefault(r.depth, []).append((r, job))
        # launch full batches, and the rest once the device runs short of work or they
        # have waited long enough
        active = sum(1 for f in self._flights if f.copies is None)
        now = time.perf_counter()
        for depth in list(self._staged):
            group = self._staged[depth]
            while len(group) >= self.max_batch:
                self._launch(group[:self.max_batch])
                group = group[self.max_batch:]
                active += 1
            if group and (active < self.min_active or now - group[0][0].admitted >= self.linger):
                self._launch(group)
                group = []
                active += 1
            if group:
                self._staged[depth] = group
            else:
                del self._staged[depth]
        self.stats.peak_inflight = max(self.stats.peak_inflight, self.inflight)

    def _upload(self, r: Request, job):
        """Runs on the thread pool: makes the job first when `job` is (capacity, footprint,
        evicted), after closing the evicted ones, then uploads the request's rows."""
        if isinstance(job, tuple):
            capacity, _, evicted = job
            self._close_all(evicted)
            try:
                job = self._job_factory(r.depth, capacity, r.precision)
            except Exception as e:  # noqa: BLE001 - reported per request
                return None, e
        try:
            job.set_points(r.points, r.normals)
            return job, None
        except Exception as e:  # noqa: BLE001 - reported per request
            return job, e

    def _batch_for(self, jobs: list):
        key = tuple(id(j) for j in jobs)
        b = self._batches.get(key)
        if b is not None:
            self._batches.move_to_end(key)
            self.stats.batches_reused += 1
            return b
        b = self._batch_factory(jobs)
        self._batches[key] = b
        if len(self._batches) > self._batch_cache:  # drop the least recent idle batches
            flying = {id(f.batch) for f in self._flights}
            for k in [k for k, v in self._batches.items() if id(v) not in flying and v is not b]:
                if len(self._batches) <= self._batch_cache:
                    break
                self._batches.pop(k).close()
        return b

    def _launch(self, group) -> None:
        reqs = [r for r, _ in group]
        jobs = [j for _, j in group]
        uniform = all(r.precision == reqs[0].precision for r in reqs) and all(
            (r.params.cycles, r.params.pre_smooth, r.params.post_smooth, r.params.point_weight)
            == (reqs[0].params.cycles, reqs[0].params.pre_smooth, reqs[0].params.post_smooth,
                reqs[0].params.point_weight) for r in reqs)
        if len(group) > 1 and uniform:
            flight = _Flight(reqs, jobs)
            try:
                flight.batch = self._batch_for(jobs)
                flight.batch.run([r.params for r in reqs])
            except Exception as e:  # noqa: BLE001
                self._fail(flight, e)
                return
            self._flights.append(flight)
            self.stats.batches += 1
        else:
            for r, j in group:
                flight = _Flight([r], [j])
                try:
                    j.run(r.params)
                except Exception as e:  # noqa: BLE001
                    self._fail(flight, e)
                    continue
                self._flights.append(flight)

    def _verify(self, r: Request) -> dict:
        """Runs on the thread pool: the verifier's mesh for `r` and the distance to r's mesh."""
        distance = self._distance
        if distance is None:
            from .metrics import mesh_distance as distance
        ref = self._verifier(r)
        p = r.points[np.isfinite(r.points).all(axis=1)]
        diag = float(np.linalg.norm(p.max(axis=0) - p.min(axis=0))) if len(p) else 0.0
        d = distance(r.mesh, ref)
        scale = diag if diag > 0.0 else 1.0
        return {"hausdorff": d.hausdorff / scale, "chamfer": d.chamfer / scale}

    def _fail(self, f: _Flight, error: Exception) -> None:
        """A flight failed: discard its jobs (and batch) and serve its requests by the
        fallback, one at a time; they are returned by the next step (completed, or with
        `error` set)."""
        if f.batch is not None:
            for key in [k for k, b in self._batches.items() if b is f.batch]:
                self._batches.pop(key)
            try:
                f.batch.close()
            except Exception:
                pass
        for j in f.jobs:
            self._discard_job(j)
        for r in f.requests:
            try:
                if self._fallback is None:
                    raise error
                r.mesh = self._fallback(r)
                r.fallback = True
                self.stats.fallbacks += 1
            except Exception as e:  # noqa: BLE001
                r.error = f"{type(error).__name__}: {error}"
                if e is not error:
                    r.error += f"; fallback {type(e).__name__}: {e}"
                r.mesh = None
                self.stats.failed += 1
            r.completed = time.perf_counter()
            self.stats.completed += 1
            self.stats.latencies.append(r.latency)
            self._ready.append(r)

    def _collect(self) -> list[Request]:
        still, finished = [], []
        for f in self._flights:
            if f.copies is None:
                try:
                    ready = f.batch.done() if f.batch is not None else f.jobs[0].done()
                    if ready:
                        results = f.batch.wait() if f.batch is not None else [f.jobs[0].wait()]
                        # host copies of the meshes on the thread pool; the flight keeps its
                        # jobs (their pinned meshes are being read) until every copy is done
                        f.copies = [self._threads.submit(mesh_from_job, j, res, len(r.points))
                                    for r, j, res in zip(f.requests, f.jobs, results)]
                except Exception as e:  # noqa: BLE001
                    self._fail(f, e)
                    continue
                still.append(f)
            elif all(c.done() for c in f.copies):
                finished.append(f)
            else:
                still.append(f)
        self._flights = still
        for f in finished:
            try:
                meshes = [c.result() for c in f.copies]
            except Exception as e:  # noqa: BLE001
                self._fail(f, e)
                continue
            now = time.perf_counter()
            for r, j, m in zip(f.requests, f.jobs, meshes):
                r.mesh = m
                r.completed = now
                self._busy.discard(id(j))
                self.stats.completed += 1
                self.stats.latencies.append(r.latency)
                if self.verify_every and r.id % self.verify_every == 0:
                    r.fp64_error = self._threads.submit(self._verify, r)
                    self.stats.verified += 1
                self._ready.append(r)
        done, self._ready = self._ready, []
        return done
