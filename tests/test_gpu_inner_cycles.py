"""GPU tests of the mixed loop's inner cycles (mpr_params.inner_cycles): the V-cycles run in
outer steps of that many FP32 V-cycles per FP64 defect r = b - A x (solver.hip run_cycles).
No reference counterpart; the bar is the mixed tolerance against the FP64 solve used
elsewhere (field within 1e-5 of its maximum, iso value to 1e-6 relative, mesh Hausdorff within
1e-4 of the bounding-box diagonal), plus bit-identity where the schedules coincide: with
cycles = 1 every inner_cycles value runs one defect and one V-cycle; 0 runs the default (1).
"""

from __future__ import annotations

import numpy as np
import pytest

import mojo_poissonrecon as mpr
import oracle
from mojo_poissonrecon import _lib, shapes
from mojo_poissonrecon.batch import Batch, reconstruct_batch
from mojo_poissonrecon.core import Job
from mojo_poissonrecon.metrics import hausdorff

pytestmark = pytest.mark.gpu


def unit(p):
    return oracle.normalize(p, 1.1)[0]


def diag_of(p):
    return float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))


@pytest.mark.parametrize("inner,cycles", [(1, 4), (2, 4), (4, 4), (2, 3), (3, 5), (8, 4)])
def test_inner_cycles_track_fp64(inner, cycles):
    p, n = shapes.torus_workload(40000, seed=11)
    kw = dict(depth=7, confidence=True, cycles=cycles)
    s64 = mpr.solve_screened_poisson(unit(p), n, **kw)
    s32 = mpr.solve_screened_poisson(unit(p), n, precision="mixed", inner_cycles=inner, **kw)
    assert np.abs(s32.field - s64.field).max() <= 1e-5 * np.abs(s64.field).max()
    assert s32.iso_value == pytest.approx(s64.iso_value, rel=1e-6)
    assert s32.residual_norm == pytest.approx(s64.residual_norm, rel=1e-3)
    m64 = mpr.reconstruct(p, n, **kw)
    m32 = mpr.reconstruct(p, n, precision="mixed", inner_cycles=inner, **kw)
    assert hausdorff(m64, m32) <= 1e-4 * diag_of(p)


def test_more_inner_cycles_stay_closer_to_fp32_rounding():
    # one FP32 V-cycle per defect tracks the FP64 iterate cycle for cycle; with all four
    # V-cycles in one outer step the FP32 rounding of the whole first correction remains
    p, n = shapes.torus_workload(60000, seed=12)
    kw = dict(depth=7, confidence=True)
    f64 = mpr.solve_screened_poisson(unit(p), n, **kw).field
    err = {k: np.abs(mpr.solve_screened_poisson(unit(p), n, precision="mixed", inner_cycles=k,
                                                **kw).field - f64).max() for k in (1, 4)}
    assert err[1] < err[4]


def test_default_is_one_and_one_cycle_runs_alike():
    p, n = shapes.torus_workload(30000, seed=13)
    kw = dict(depth=7, confidence=True, precision="mixed")
    a = mpr.solve_screened_poisson(unit(p), n, inner_cycles=0, **kw)
    b = mpr.solve_screened_poisson(unit(p), n, inner_cycles=1, **kw)
    assert np.array_equal(a.field, b.field) and a.residual_norm == b.residual_norm
    one = [mpr.solve_screened_poisson(unit(p), n, cycles=1, inner_cycles=k, **kw)
           for k in (1, 2, 5)]
    for s in one[1:]:
        assert np.array_equal(s.field, one[0].field)
        assert s.residual_norm == one[0].residual_norm


def test_batch_with_inner_cycles_equals_solo_runs():
    makers = (shapes.sphere, shapes.torus, shapes.superquadric)
    clouds = [makers[i % 3](3000 + 2000 * i, seed=20 + i) for i in range(5)]
    depths = [5, 8, 6, 7, 8]
    for inner in (1, 3):
        batch = reconstruct_batch(clouds, depths=depths, precision="mixed", inner_cycles=inner)
        for (p, n), d, m in zip(clouds, depths, batch):
            solo = mpr.reconstruct(p, n, depth=d, precision="mixed", inner_cycles=inner)
            assert m.iso_value == solo.iso_value and m.residual_norm == solo.residual_norm
            assert np.array_equal(m.vertices, solo.vertices)
            assert np.array_equal(m.faces, solo.faces)


def test_batch_rejects_mixed_inner_cycles():
    p, n = shapes.sphere(3000, seed=1)
    jobs = [Job(5, 3000, _lib.PRECISION_MIXED) for _ in range(2)]
    try:
        for j in jobs:
            j.set_points(unit(p), n)
        b = Batch(jobs)
        try:
            with pytest.raises(_lib.DeviceError, match="inner cycles"):
                b.run([Job.params(_lib.STAGE_SOLVE, inner_cycles=1),
                       Job.params(_lib.STAGE_SOLVE, inner_cycles=2)])
        finally:
            b.close()
    finally:
        for j in jobs:
            j.close()
