"""GPU tests of the fused smoothing (march.hip): one or two red-black iterations of an FP32
level in one z-marching pass must give the single sweeps' results bit for bit.

The fused pass repeats relax_node's arithmetic per node (capi.mojo:309-410) in the same
order, and every value it reads was produced by the same sweep sequence, so the mixed
solve's fields, iso value, residual norm and mesh are identical for every setting of
mpr_tune_set_march (0: single sweeps, 1 or 2 iterations fused; levels of 257^3 march): one
request at depths 5-9 (depth 8's finest level marches from the defect pass's first sweep,
depth 9's second level from zero; depths 5-7 have no such level and must be untouched), the
parameter corners the schedule has its own code for (no screening, one cycle, six cycles,
heavy screening), an uneven batch of depths 5-8, and the smoothing counts that reach every
branch of the schedule. Its division by an unscreened diagonal (one multiplication and one
FMA with the reciprocal split into two floats) must equal the IEEE quotient for every finite
float numerator, on the levels of 65^3 to 1025^3. Tolerance: none (array_equal, bit-pattern equality).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import mojo_poissonrecon as mpr
from mojo_poissonrecon import _lib, shapes, tune
from mojo_poissonrecon.batch import reconstruct_batch
from mojo_poissonrecon.core import Job

pytestmark = pytest.mark.gpu

SETTINGS = (0, 1, 2)


@pytest.fixture
def march_default():
    """Restores the default and drops pooled jobs built under another setting."""
    default = tune.march_iterations()
    yield
    tune.set_march_iterations(default)
    mpr.release_device_memory()


def unit_cube(points):
    lo, hi = points.min(axis=0), points.max(axis=0)
    return (points - (lo + hi) / 2) / (float((hi - lo).max()) * 1.1) + 0.5


def run(points, normals, iterations, **kw):
    tune.set_march_iterations(iterations)
    mpr.release_device_memory()  # pooled jobs keep the schedule they were built with
    sol = mpr.solve_screened_poisson(unit_cube(points), normals, precision="mixed", **kw)
    mesh = mpr.reconstruct(points, normals, precision="mixed", **kw)
    return sol, mesh


def assert_identical(a, b):
    (sa, ma), (sb, mb) = a, b
    assert np.array_equal(sa.field, sb.field)
    assert np.array_equal(sa.rhs, sb.rhs)
    assert sa.iso_value == sb.iso_value
    assert sa.residual_norm == sb.residual_norm
    assert np.array_equal(ma.vertices, mb.vertices)
    assert np.array_equal(ma.faces, mb.faces)
    assert np.array_equal(ma.density, mb.density)


@pytest.mark.parametrize("depth,count",
                         [(5, 3000), (6, 12000), (7, 40000), (8, 200000), (9, 300000)])
def test_fused_smoothing_bit_identical(march_default, depth, count):
    if depth == 8:
        p, n = shapes.torus_workload(count, seed=depth)
    else:
        p, n = shapes.perturb(*shapes.torus(count, seed=depth), seed=depth, sigma=0.002,
                              degrees=5.0)
    kw = dict(depth=depth, confidence=True)
    base = run(p, n, 0, **kw)
    for it in SETTINGS[1:]:
        assert_identical(base, run(p, n, it, **kw))


@pytest.mark.parametrize("kw", [dict(point_weight=0.0), dict(cycles=1), dict(cycles=6),
                                dict(point_weight=40.0, samples_per_node=3.0)])
def test_fused_smoothing_parameters(march_default, kw):
    p, n = shapes.superquadric(60000, seed=3)
    kw = dict(depth=8, **kw)
    base = run(p, n, 0, **kw)
    for it in SETTINGS[1:]:
        assert_identical(base, run(p, n, it, **kw))


def test_fused_smoothing_uneven_batch(march_default):
    makers = (shapes.sphere, shapes.torus, shapes.superquadric)
    clouds = [makers[i % 3](2000 + 3000 * i, seed=10 + i) for i in range(6)]
    depths = [5, 8, 6, 7, 8, 6]
    out = []
    for it in SETTINGS:
        tune.set_march_iterations(it)
        out.append(reconstruct_batch(clouds, depths=depths, precision="mixed"))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert a.iso_value == b.iso_value
            assert a.residual_norm == b.residual_norm
            assert np.array_equal(a.vertices, b.vertices)
            assert np.array_equal(a.faces, b.faces)


@pytest.mark.parametrize("pre,post", [(3, 2), (2, 3), (2, 2), (1, 1), (4, 4), (3, 1), (1, 3)])
def test_fused_smoothing_smoothing_counts(march_default, pre, post):
    # every branch of the schedule: the pass's own correction output (post = fused
    # iterations), a pre-smoothing too short to fuse after the defect pass's first sweep,
    # one-iteration phases, longer ones
    p, n = shapes.torus(60000, seed=5)
    p, n = shapes.perturb(p, n, seed=5, sigma=0.002, degrees=5.0)
    out = []
    for it in SETTINGS:
        tune.set_march_iterations(it)
        job = Job(8, len(p), _lib.PRECISION_MIXED)
        try:
            job.set_points(p, n)
            job.run(Job.params(_lib.STAGE_NORMALIZE | _lib.STAGE_SOLVE | _lib.STAGE_GRIDS, 4.0,
                               1.5, confidence=True, pre_smooth=pre, post_smooth=post))
            res = job.wait()
            out.append((res.iso_value, res.residual_norm, job.fetch_grids()[0]))
        finally:
            job.close()
    for iso, rn, field in out[1:]:
        assert iso == out[0][0] and rn == out[0][1]
        assert np.array_equal(field, out[0][2])


@pytest.mark.parametrize("n", [65, 257, 513, 1025])
def test_fast_division_exact_for_every_float(n):
    # 4 diagonals x 2^32 numerators each, on the device
    assert tune.fast_division_mismatches(n) == 0


def test_fused_smoothing_timing_hook(march_default):
    tune.set_march_iterations(2)
    p, n = shapes.torus_workload(200000, seed=1)
    job = Job(8, len(p), _lib.PRECISION_MIXED)
    try:
        job.set_points(p, n)
        job.run(Job.params(_lib.STAGE_NORMALIZE | _lib.STAGE_SOLVE, 4.0, 1.5, confidence=True))
        job.wait()
        ms, nbytes = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load_library().mpr_job_time_kernel(job.handle, 14, 3, ctypes.byref(ms),
                                                           ctypes.byref(nbytes)),
                   "mpr_job_time_kernel")
        assert ms.value > 0 and nbytes.value > 2.5 * 4 * 257**3
    finally:
        job.close()
