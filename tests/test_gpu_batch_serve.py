"""GPU tests of the extensions: mixed precision, batched reconstruction, graph replay of a
batch, and the continuous-batching server (each result the same alone and in the mix).

Tolerances. Every stage is deterministic and an instance's arithmetic does not depend on
what it runs with, so "the same alone and in the mix" (batches, graph replays, the server)
is bit-identical: iso value, residual norm, vertices, faces and vertex density. Mixed
precision against FP64: iso within 1e-6 relative, field within 1e-5 of its maximum, and
meshes within a Hausdorff bound (vertex to nearest vertex, both directions) of 1e-4 of the
bounding-box diagonal with vertex counts within 2.
"""

from __future__ import annotations

import numpy as np
import pytest

import mojo_poissonrecon as mpr
import oracle
from mojo_poissonrecon import shapes
from mojo_poissonrecon.batch import Batch, reconstruct_batch
from mojo_poissonrecon.core import Job, prepare_reconstruct
from mojo_poissonrecon.serve import ReconstructionServer, job_capacity

pytestmark = pytest.mark.gpu


def hausdorff_bound(a, b):
    from scipy.spatial import cKDTree
    if not len(a) or not len(b):
        return 0.0 if len(a) == len(b) else np.inf
    return max(cKDTree(b).query(a)[0].max(), cKDTree(a).query(b)[0].max())


def same_mesh(m1, m2, diag, tol=1e-9):
    assert abs(len(m1.vertices) - len(m2.vertices)) <= 2
    assert hausdorff_bound(m1.vertices, m2.vertices) <= tol * diag


def identical_mesh(m1, m2):
    # every stage is deterministic (the splat sums in point order, reductions in a fixed
    # order) and an instance's arithmetic does not depend on its batch: bit-identical
    assert m1.iso_value == m2.iso_value
    assert m1.residual_norm == m2.residual_norm
    assert np.array_equal(m1.vertices, m2.vertices)
    assert np.array_equal(m1.faces, m2.faces)
    assert np.array_equal(m1.density, m2.density)


def clouds(n):
    out = []
    makers = (shapes.sphere, shapes.torus, shapes.superquadric)
    for i in range(n):
        p, q = makers[i % 3](1500 + 700 * i, seed=i)
        p, q = shapes.perturb(p, q, seed=i, sigma=0.001 * (i % 4), degrees=2.0)
        out.append((p, q))
    return out


def diag_of(p):
    return float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))


def test_mixed_precision_tracks_fp64():
    p, n = shapes.torus_workload(30000, seed=3)
    for depth in (5, 7):
        m64 = mpr.reconstruct(p, n, depth=depth, confidence=True, precision="fp64")
        m32 = mpr.reconstruct(p, n, depth=depth, confidence=True, precision="mixed")
        assert m32.iso_value == pytest.approx(m64.iso_value, rel=1e-6)
        same_mesh(m32, m64, diag_of(p), tol=1e-4)
        s64 = mpr.solve_screened_poisson(*oracle.normalize(p, 1.1)[:1], n, depth=depth,
                                         confidence=True)
        s32 = mpr.solve_screened_poisson(*oracle.normalize(p, 1.1)[:1], n, depth=depth,
                                         confidence=True, precision="mixed")
        assert np.abs(s32.field - s64.field).max() <= 1e-5 * np.abs(s64.field).max()
        assert s32.residual_norm == pytest.approx(s64.residual_norm, rel=1e-3)


@pytest.mark.parametrize("point_weight,cycles", [(0.0, 4), (25.0, 2), (4.0, 1)])
def test_mixed_precision_tracks_fp64_across_parameters(point_weight, cycles):
    # no screening, strong screening and a single cycle: the mixed field stays within the
    # same bounds of the FP64 field
    p, n = shapes.torus_workload(20000, seed=5)
    unit = oracle.normalize(p, 1.1)[0]
    kw = dict(depth=6, confidence=True, point_weight=point_weight, cycles=cycles)
    s64 = mpr.solve_screened_poisson(unit, n, **kw)
    s32 = mpr.solve_screened_poisson(unit, n, precision="mixed", **kw)
    assert s32.iso_value == pytest.approx(s64.iso_value, rel=1e-6)
    assert np.abs(s32.field - s64.field).max() <= 1e-5 * np.abs(s64.field).max()
    assert s32.residual_norm == pytest.approx(s64.residual_norm, rel=1e-3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_solo_runs(precision):
    cs = clouds(7)
    depths = [2, 4, 5, 6, 7, 5, 6]
    batch = reconstruct_batch(cs, depths=depths, precision=precision)
    for (p, n), d, m in zip(cs, depths, batch):
        solo = mpr.reconstruct(p, n, depth=d, precision=precision)
        identical_mesh(m, solo)


def test_batch_fp64_matches_oracle_and_handles_empty_clouds():
    p, n = shapes.sphere(2000, 4)
    empty = (np.zeros((3, 3)), np.zeros((3, 3)))  # zero normals: no valid rows
    out = reconstruct_batch([(p, n), empty, (p, n)], depths=[5, 5, 6], precision="fp64")
    want = oracle.reconstruct(p, n, depth=5)
    assert out[0].iso_value == pytest.approx(want["iso_value"], rel=1e-10)
    assert np.array_equal(out[0].faces, want["faces"])
    assert out[1].is_empty
    assert not out[2].is_empty


def test_mixed_batch_with_an_empty_job_matches_solo_runs():
    # a job with no points keeps the filled solution (0) while the others in the batch start
    # from the constant initial guess without a fill; each result equals its solo run
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import mesh_from_job
    p, n = shapes.torus(3000, seed=6)
    reqs = [prepare_reconstruct(p, n, depth=d, precision="mixed") for d in (5, 6)]

    def solo(r):
        j = Job(r.depth, len(r.points), r.precision)
        try:
            j.set_points(r.points, r.normals)
            j.run(r.params)
            return mesh_from_job(j, j.wait(), len(r.points))
        finally:
            j.close()

    want = [solo(r) for r in reqs]
    jobs = [Job(d, 3000, _lib.PRECISION_MIXED) for d in (5, 5, 6)]
    try:
        jobs[0].set_points(reqs[0].points, reqs[0].normals)
        jobs[1].set_points(np.zeros((0, 3)), np.zeros((0, 3)))
        jobs[2].set_points(reqs[1].points, reqs[1].normals)
        b = Batch(jobs)
        try:
            b.run([reqs[0].params, reqs[0].params, reqs[1].params])
            res = b.wait()
        finally:
            b.close()
        assert res[1].vertex_count == 0 and res[1].face_count == 0
        identical_mesh(mesh_from_job(jobs[0], res[0], len(reqs[0].points)), want[0])
        identical_mesh(mesh_from_job(jobs[2], res[2], len(reqs[1].points)), want[1])
    finally:
        for j in jobs:
            j.close()


def test_batch_graph_replays_match_eager():
    cs = clouds(5)
    depths = [3, 5, 6, 7, 4]
    reqs = [prepare_reconstruct(p, n, depth=d, precision="mixed") for (p, n), d in zip(cs, depths)]
    jobs = [Job(r.depth, len(r.points), r.precision) for r in reqs]
    for j, r in zip(jobs, reqs):
        j.set_points(r.points, r.normals)
    b = Batch(jobs)
    b.run([r.params for r in reqs])
    eager = [(res.iso_value, j.fetch_mesh(res)) for j, res in zip(jobs, b.wait())]
    b.capture([r.params for r in reqs])
    for _ in range(3):
        b.launch()
        for (iso, (v, f, d)), j, res in zip(eager, jobs, b.wait()):
            assert res.iso_value == iso
            v2, f2, d2 = j.fetch_mesh(res)
            assert np.array_equal(v2, v) and np.array_equal(f2, f) and np.array_equal(d2, d)
    b.close()
    for j in jobs:
        j.close()


def test_server_results_equal_solo_runs():
    rng = np.random.default_rng(0)
    server = ReconstructionServer(max_inflight=6, max_batch=3, precision="mixed")
    cs = clouds(10)
    depths = [int(rng.choice([5, 6, 7])) for _ in cs]
    reqs = {}
    done = []
    for i, ((p, n), d) in enumerate(zip(cs, depths)):
        r = server.submit(p, n, depth=d)
        reqs[r.id] = (p, n, d)
        if i % 4 == 3:  # bursts of arrivals: admitted together as a batch
            done += server.step()
    done += server.drain()
    server.close()
    assert len(done) == len(cs)
    assert server.stats.batches >= 1
    for r in done:
        p, n, d = reqs[r.id]
        solo = mpr.reconstruct(p, n, depth=d, precision="mixed")
        identical_mesh(r.mesh, solo)
        assert r.latency > 0


def test_server_verifies_sampled_results_against_fp64():
    # verify_every: the verifier (plain FP64 reconstruct) re-solves every 2nd result on the
    # host threads; the mixed meshes lie within 1e-4 of the diagonal of the FP64 ones
    rng = np.random.default_rng(3)
    server = ReconstructionServer(max_inflight=6, max_batch=3, precision="mixed",
                                  verify_every=2)
    cs = clouds(6)
    reqs = [server.submit(p, n, depth=int(rng.choice([5, 6, 7]))) for p, n in cs]
    done = server.drain()
    errors = {r.id: r.fp64_error.result() for r in reqs if r.fp64_error is not None}
    server.close()
    assert len(done) == 6 and server.stats.verified == 3
    assert sorted(errors) == [r.id for r in reqs if r.id % 2 == 0]
    for e in errors.values():
        assert 0.0 <= e["chamfer"] <= e["hausdorff"] <= 1e-4


def test_server_admission_respects_memory_budget():
    p, n = shapes.sphere(3000, 1)
    fp = Job.footprint(6, job_capacity(len(p)), 1)
    server = ReconstructionServer(hbm_budget=int(2.5 * fp), max_inflight=8, max_batch=1,
                                  precision="mixed")
    for i in range(6):
        server.submit(p, n, depth=6)
    server.step()
    assert server.inflight <= 2
    done = server.drain()
    assert len(done) == 6 and server.stats.peak_bytes <= 2.5 * fp
    server.close()


def test_launch_shapes_do_not_change_results():
    from mojo_poissonrecon import tune
    p, n = shapes.torus_workload(20000, seed=5)
    keys = [(k, t, s) for k in tune.KERNELS for t in ("fp32", "fp64") for s in range(5, 8)]
    saved = {key: (tune.tuned_block(*key), tune.tuned_groups(*key), tune.tuned_planes(*key))
             for key in keys}
    try:
        base = {prec: mpr.reconstruct(p, n, depth=7, precision=prec) for prec in ("fp64", "mixed")}
        # (block, groups, planes): planes > 1 is the sweep's z-column form; 129 = 4 * 32 + 1
        # and 65 = 4 * 16 + 1 nodes per side leave a last column of one plane
        for block, groups, planes in ((64, 2, 1), (1024, 1, 1), (128, 2, 1), (256, 1, 2),
                                      (64, 1, 4), (1024, 1, 4)):
            for key in keys:
                tune.set_tuned_block(*key, block)
                if key[0] == "relax":
                    tune.set_tuned_groups(*key, groups)
                    tune.set_tuned_planes(*key, planes)
            for prec, m in base.items():
                identical_mesh(mpr.reconstruct(p, n, depth=7, precision=prec), m)
        chosen = tune.autotune(max_depth=7, reps=3)
        assert {b for b, _, _ in chosen.values()} <= set(tune.CANDIDATES)
        assert {g for _, g, _ in chosen.values()} <= set(tune.GROUP_CANDIDATES)
        assert {z for _, _, z in chosen.values()} <= set(tune.PLANE_CANDIDATES)
        assert all(g == 1 or z == 1 for _, g, z in chosen.values())
        assert len(chosen) == 3 * 2 + 3 * 2 + 3  # relax, residual x2 precisions; defect fp64
        for prec, m in base.items():
            identical_mesh(mpr.reconstruct(p, n, depth=7, precision=prec), m)
    finally:
        for key, (b, g, z) in saved.items():
            tune.set_tuned_block(*key, b)
            if key[0] == "relax":
                tune.set_tuned_groups(*key, g)
                tune.set_tuned_planes(*key, z)
