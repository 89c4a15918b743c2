"""CPU tests of the serving loop's failure handling, with stand-in jobs and batches (no
device): a flight whose batch fails at launch, at the completion query or at the wait is
isolated - its jobs are discarded, its requests are served by the fallback (or completed
with their error when the fallback fails too) - and the server keeps serving later
requests, then closes cleanly."""

from __future__ import annotations

import numpy as np
import pytest

from mojo_poissonrecon import _lib
from mojo_poissonrecon.core import Mesh
from mojo_poissonrecon.serve import ReconstructionServer


class FakeJob:
    def __init__(self, depth, capacity, precision):
        self.depth, self.max_points, self.precision = depth, capacity, precision
        self.count = 0
        self.closed = False
        self.runs = 0

    def set_points(self, points, normals):
        self.count = len(points)

    def run(self, params):
        self.runs += 1

    def done(self):
        return True

    def wait(self):
        r = _lib.mpr_result()
        r.accepted = r.valid_rows = self.count
        r.vertex_count, r.face_count = 3, 1
        return r

    def fetch_mesh(self, result):
        return np.zeros((3, 3)), np.array([[0, 1, 2]]), np.zeros(3)

    def close(self):
        self.closed = True


class FakeBatch:
    fail_at = {}  # batch number -> "run" | "done" | "wait"
    made = 0

    def __init__(self, jobs):
        self.jobs = jobs
        self.number = FakeBatch.made
        FakeBatch.made += 1
        self.closed = False

    def _maybe(self, where):
        if FakeBatch.fail_at.get(self.number) == where:
            raise _lib.DeviceError(f"batch {self.number} failed at {where}")

    def run(self, params):
        self._maybe("run")
        for j in self.jobs:
            j.runs += 1

    def done(self):
        self._maybe("done")
        return True

    def wait(self):
        self._maybe("wait")
        return [j.wait() for j in self.jobs]

    def close(self):
        self.closed = True


def fallback_mesh(r):
    return Mesh(np.ones((1, 3)), np.empty((0, 3), np.int64), np.ones(1), 0.25, 0.0)


def make_server(fallback=fallback_mesh, jobs=None):
    made = [] if jobs is None else jobs

    def job_factory(d, c, p):
        made.append(FakeJob(d, c, p))
        return made[-1]

    return ReconstructionServer(hbm_budget=1 << 50, max_inflight=6, max_batch=3,
                                job_factory=job_factory, batch_factory=FakeBatch,
                                fallback=fallback), made


def cloud(n=100):
    rng = np.random.default_rng(n)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3))


@pytest.mark.parametrize("where", ["run", "done", "wait"])
def test_a_failed_batch_is_isolated_and_served_by_the_fallback(where):
    FakeBatch.made = 0
    FakeBatch.fail_at = {0: where}
    server, jobs = make_server()
    first = [server.submit(*cloud(), depth=6) for _ in range(3)]
    done = server.step() + server.drain()
    assert {r.id for r in done} == {r.id for r in first}
    assert all(r.fallback and r.error is None and r.mesh.iso_value == 0.25 for r in done)
    assert server.stats.fallbacks == 3 and server.stats.failed == 0
    assert all(j.closed for j in jobs[:3])  # the failed flight's jobs are discarded
    # the server keeps serving: a new batch on new jobs, not the fallback
    later = [server.submit(*cloud(50 + k), depth=6) for k in range(3)]
    done = server.drain()
    assert {r.id for r in done} == {r.id for r in later}
    assert not any(r.fallback for r in done) and all(r.mesh.faces.shape == (1, 3) for r in done)
    assert not any(j.closed for j in jobs[3:])
    server.close()
    assert all(j.closed for j in jobs)


def test_a_failed_fallback_completes_the_request_with_its_error():
    FakeBatch.made = 0
    FakeBatch.fail_at = {0: "wait"}

    def broken(r):
        raise RuntimeError("no device")

    server, _ = make_server(fallback=broken)
    for _ in range(3):
        server.submit(*cloud(), depth=7)
    done = server.drain()
    assert len(done) == 3 and server.stats.failed == 3
    assert all(r.mesh is None and "batch 0 failed at wait" in r.error and "no device" in r.error
               for r in done)
    server.close()


def test_batches_over_the_same_jobs_are_reused():
    FakeBatch.made = 0
    FakeBatch.fail_at = {}
    server, jobs = make_server()
    for rnd in range(3):
        for _ in range(3):
            server.submit(*cloud(), depth=6)
        assert len(server.drain()) == 3
    assert len(jobs) == 3 and FakeBatch.made == 1
    assert server.stats.batches == 3 and server.stats.batches_reused == 2
    server.close()


def test_a_failed_upload_is_isolated_and_the_rest_of_its_tick_runs():
    # uploads run on the thread pool without blocking the loop; the requests admitted in
    # one tick launch together once all their uploads are done, minus those that failed
    FakeBatch.made = 0
    FakeBatch.fail_at = {}
    made = []

    def job_factory(d, c, p):
        j = FakeJob(d, c, p)
        if not made:
            def broken(points, normals):
                raise RuntimeError("upload failed")
            j.set_points = broken
        made.append(j)
        return j

    server = ReconstructionServer(hbm_budget=1 << 50, max_inflight=6, max_batch=3,
                                  job_factory=job_factory, batch_factory=FakeBatch,
                                  fallback=fallback_mesh)
    reqs = [server.submit(*cloud(), depth=6) for _ in range(3)]
    done = server.drain()
    assert {r.id for r in done} == {r.id for r in reqs}
    assert [r.id for r in done if r.fallback] == [reqs[0].id]
    assert made[0].closed and not any(j.closed for j in made[1:])
    assert FakeBatch.made == 1 and server.stats.batches == 1  # the other two, as one batch
    server.close()


def test_a_failed_job_creation_is_isolated_and_its_bytes_are_returned():
    # new pool jobs are made on the thread pool; one whose creation fails is served by the
    # fallback, the bytes it held are given back, and the rest of its tick runs
    FakeBatch.made = 0
    FakeBatch.fail_at = {}
    made = []

    def job_factory(d, c, p):
        if not made:
            made.append(None)
            raise _lib.DeviceError("out of memory")
        made.append(FakeJob(d, c, p))
        return made[-1]

    server = ReconstructionServer(hbm_budget=1 << 50, max_inflight=6, max_batch=3,
                                  job_factory=job_factory, batch_factory=FakeBatch,
                                  fallback=fallback_mesh)
    reqs = [server.submit(*cloud(), depth=6) for _ in range(3)]
    done = server.drain()
    assert {r.id for r in done} == {r.id for r in reqs}
    assert sum(r.fallback for r in done) == 1 and server.stats.fallbacks == 1
    assert server.stats.batches == 1 and len(made) == 3
    pooled = sum(server._footprint(j) for j in made[1:])
    assert server._bytes == pooled  # the failed creation's reservation was returned
    server.close()
    assert all(j.closed for j in made[1:])


def test_evicted_jobs_are_closed_off_the_loop_and_replaced():
    # a budget of two depth-6 jobs: depth-7 requests evict the idle depth-6 jobs, whose
    # close runs on the thread pool before the new jobs are made
    FakeBatch.made = 0
    FakeBatch.fail_at = {}
    made = []

    def job_factory(d, c, p):
        made.append(FakeJob(d, c, p))
        return made[-1]

    from mojo_poissonrecon.core import Job
    from mojo_poissonrecon.serve import job_capacity
    fp7 = Job.footprint(7, job_capacity(100, 7, _lib.PRECISION_MIXED), _lib.PRECISION_MIXED)
    server = ReconstructionServer(hbm_budget=fp7, max_inflight=6, max_batch=3,
                                  job_factory=job_factory, batch_factory=FakeBatch,
                                  fallback=fallback_mesh)
    for _ in range(2):
        server.submit(*cloud(), depth=6)
    assert len(server.drain()) == 2
    six = list(made)
    assert len(six) == 2 and not any(j.closed for j in six)
    r = server.submit(*cloud(), depth=7)
    done = server.drain()
    assert [x.id for x in done] == [r.id] and not done[0].fallback
    assert all(j.closed for j in six) and len(made) == 3 and made[2].depth == 7
    assert server._bytes == fp7
    server.close()


def test_verify_every_sends_every_kth_result_to_the_verifier():
    # the verifier (the FP64 solo path by default) re-solves every k-th completed request on
    # the thread pool; its future gives the distances over the bbox diagonal
    from types import SimpleNamespace
    FakeBatch.made = 0
    FakeBatch.fail_at = {}
    verified = []

    def verifier(r):
        verified.append(r.id)
        return Mesh(np.zeros((3, 3)), np.array([[0, 1, 2]]), np.zeros(3), 0.5, 0.0)

    def distance(a, b):
        return SimpleNamespace(hausdorff=2.0, chamfer=1.0)

    server = ReconstructionServer(hbm_budget=1 << 50, max_inflight=6, max_batch=3,
                                  job_factory=lambda d, c, p: FakeJob(d, c, p),
                                  batch_factory=FakeBatch, fallback=fallback_mesh,
                                  verify_every=2, verifier=verifier, distance=distance)
    points = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0], [np.nan, 0.0, 0.0]] * 10)
    normals = np.ones_like(points)
    reqs = [server.submit(points, normals, depth=6) for _ in range(5)]
    done = server.drain()
    server.close()
    assert len(done) == 5 and server.stats.verified == 3
    for r in reqs:
        if r.id % 2 == 0:
            err = r.fp64_error.result()
            assert err == {"hausdorff": 2.0 / 5.0, "chamfer": 1.0 / 5.0}  # finite rows: diag 5
        else:
            assert r.fp64_error is None
    assert sorted(verified) == [r.id for r in reqs if r.id % 2 == 0]
    with pytest.raises(ValueError):
        ReconstructionServer(hbm_budget=1 << 50, verify_every=-1)


def test_pooled_capacity_grows_while_the_footprint_barely_does():
    from mojo_poissonrecon.core import Job
    from mojo_poissonrecon.serve import job_capacity
    mixed = _lib.PRECISION_MIXED
    assert job_capacity(5000) == 16384 and job_capacity(300_000) == 1 << 19
    for depth in (5, 6, 7, 8, 9):
        floor = Job.footprint(depth, 16384, mixed)
        for points in (100, 20_000, 300_000, 2_000_000):
            c = job_capacity(points, depth, mixed)
            assert c >= job_capacity(points) and c >= points and c & (c - 1) == 0
            grown = c > job_capacity(points)
            assert not grown or Job.footprint(depth, c, mixed) <= 1.03 * floor
            if c < 1 << 24:  # the next class would pass the 3 % bound
                assert Job.footprint(depth, 2 * c, mixed) > 1.03 * floor
    # a depth-9 job serves every cloud of up to 512k points; shallow jobs stay small
    assert job_capacity(5000, 9, mixed) >= 1 << 19
    assert job_capacity(5000, 6, mixed) == 16384
