"""GPU parity of the device-resident pipeline (public API) against the oracle.

solve_screened_poisson, extract_isosurface, build_octree, reconstruct and reconstruct_mesh
run through libmpr_hip.so and are compared with the oracle's restatement of
core.py + capi.mojo on the same inputs. Tolerances: the splat's FP64 atomics add in a
different order than the reference's serial loop, so grids agree to ~1e-13 relative and
the solution to ~1e-11 relative; the octree arrays are bit-exact; extraction on an
identical field is bit-exact.
"""

from __future__ import annotations

import numpy as np
import pytest

import mojo_poissonrecon as mpr
import oracle
from mojo_poissonrecon import shapes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("depth,count,confidence", [(4, 500, False), (5, 2000, True),
                                                    (6, 8000, False), (7, 20000, True)])
def test_solve_matches_oracle(depth, count, confidence):
    p, n = shapes.perturb(*shapes.sphere(count, depth), seed=depth, sigma=0.002, degrees=5.0,
                          length_range=(0.5, 1.5))
    unit, _, _ = oracle.normalize(p, 1.1)
    got = mpr.solve_screened_poisson(unit, n, depth=depth, confidence=confidence)
    want = oracle.solve(unit, n, depth, confidence=confidence)
    # the splat sums in the reference's point order and every later stage repeats its
    # arithmetic, so the grids are bit-identical; the two global reductions (iso value and
    # residual norm) add in another order
    assert np.array_equal(got.density, want["density"])
    assert np.array_equal(got.screen, want["screen"])
    assert np.array_equal(got.rhs, want["rhs"])
    assert np.array_equal(got.field, want["field"])
    assert got.iso_value == pytest.approx(want["iso_value"], rel=1e-12)
    assert got.residual_norm == pytest.approx(want["residual_norm"], rel=1e-12)
    tree = oracle.build_octree(unit, depth)
    for name in ("depth", "morton", "parent", "child_start", "child_count", "point_leaf"):
        assert np.array_equal(getattr(got.octree, name), tree[name]), name


@pytest.mark.parametrize("point_weight,samples_per_node,cycles",
                         [(0.0, 1.5, 4), (25.0, 1.0, 2), (4.0, 3.0, 1), (1e-3, 1.5, 5)])
def test_solve_parameters_match_oracle(point_weight, samples_per_node, cycles):
    # no screening (pure Poisson: the sweeps' diagonal is the degree alone), strong
    # screening, other samples per node and cycle counts: grids bit-identical to the oracle
    p, n = shapes.perturb(*shapes.torus(6000, 3), seed=3, sigma=0.002, degrees=5.0,
                          length_range=(0.5, 1.5))
    unit, _, _ = oracle.normalize(p, 1.1)
    got = mpr.solve_screened_poisson(unit, n, depth=6, point_weight=point_weight,
                                     samples_per_node=samples_per_node, cycles=cycles,
                                     confidence=True)
    want = oracle.solve(unit, n, 6, point_weight=point_weight, confidence=True, cycles=cycles,
                        samples_per_node=samples_per_node)
    assert np.array_equal(got.density, want["density"])
    assert np.array_equal(got.screen, want["screen"])
    assert np.array_equal(got.rhs, want["rhs"])
    assert np.array_equal(got.field, want["field"])
    assert got.iso_value == pytest.approx(want["iso_value"], rel=1e-12)
    assert got.residual_norm == pytest.approx(want["residual_norm"], rel=1e-12)


def test_solve_of_isolated_points_is_bit_exact():
    # points in distinct cells far apart (every node receives one point's contributions):
    # the whole solve is the reference's arithmetic and must match the oracle bit for bit
    g = np.linspace(0.1, 0.9, 4)
    unit = np.array([[x, y, z] for x in g for y in g for z in g])
    nrm = unit - 0.5
    nrm[np.all(nrm == 0, axis=1)] = 1.0
    got = mpr.solve_screened_poisson(unit, nrm, depth=5)
    want = oracle.solve(unit, nrm, 5)
    assert np.array_equal(got.density, want["density"])
    assert np.array_equal(got.rhs, want["rhs"])
    assert np.array_equal(got.field, want["field"])


def test_solve_with_boundary_and_duplicate_points_is_bit_exact():
    # points on the faces, edges and corners of the unit cube (the splat clamps the cell
    # index there, capi.mojo:127-173) and exact duplicates (equal node keys, summed in point
    # order by the stable sort): the grids equal the oracle's bit for bit
    rng = np.random.default_rng(17)
    corners = np.array([[x, y, z] for x in (0.0, 1.0) for y in (0.0, 1.0) for z in (0.0, 1.0)])
    faces = rng.random((300, 3))
    faces[np.arange(300), rng.integers(0, 3, 300)] = rng.integers(0, 2, 300).astype(float)
    inner = rng.random((400, 3)) * 0.8 + 0.1
    unit = np.concatenate([corners, faces, inner, inner[:50], faces[:20]])
    nrm = rng.normal(size=unit.shape)
    for depth in (3, 5):
        got = mpr.solve_screened_poisson(unit, nrm, depth=depth, confidence=True)
        want = oracle.solve(unit, nrm, depth, confidence=True)
        assert np.array_equal(got.density, want["density"])
        assert np.array_equal(got.screen, want["screen"])
        assert np.array_equal(got.rhs, want["rhs"])
        assert np.array_equal(got.field, want["field"])


def test_confidence_known_answers():
    # reference tests/test_kernels.py:46-58
    points = np.array([[0.25, 0.25, 0.25], [0.75, 0.75, 0.75]])
    normals = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    plain = mpr.solve_screened_poisson(points, normals, depth=2, cycles=1, confidence=False,
                                       samples_per_node=1.0)
    weighted = mpr.solve_screened_poisson(points, normals, depth=2, cycles=1, confidence=True,
                                          samples_per_node=1.0)
    assert plain.density.sum() == 4.0
    assert weighted.density.sum() == 6.0
    assert plain.screen.sum() == 2.0
    assert weighted.screen.sum() == 3.0


def test_octree_reference_case_and_large():
    rng = np.random.default_rng(3)
    points = rng.random((200, 3))
    tree = mpr.build_octree(points, 7)
    want = oracle.build_octree(points, 7)
    for name in want:
        assert np.array_equal(getattr(tree, name), want[name]), name
    big = np.random.default_rng(9).random((300_000, 3))
    big[:5] = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]]
    tree = mpr.build_octree(big, 9)
    want = oracle.build_octree(big, 9)
    for name in want:
        assert np.array_equal(getattr(tree, name), want[name]), name
    empty = mpr.build_octree(np.empty((0, 3)), 4)
    assert empty.node_count == 1 and empty.parent[0] == -1 and empty.point_leaf.size == 0


@pytest.mark.parametrize("n", [2, 13, 40, 97])
def test_extract_isosurface_bit_exact_on_identical_field(n):
    rng = np.random.default_rng(n)
    ax = np.linspace(0, 1, n)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    field = (x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.55) ** 2 - 0.3 ** 2
    field += 0.01 * rng.normal(size=field.shape)
    density = rng.random(field.shape)
    mesh = mpr.extract_isosurface(field, iso_value=0.001, density=density)
    v, f, d = oracle.extract_isosurface(field, 0.001, density)
    assert np.array_equal(mesh.vertices, v)
    assert np.array_equal(mesh.faces, f)
    assert np.array_equal(mesh.density, d)


@pytest.mark.parametrize("n", [3, 32, 33, 34, 65, 66])
def test_extract_random_field_bit_exact(n):
    # dense sign changes around every 32-cell run boundary of the device's bit masks
    rng = np.random.default_rng(1000 + n)
    field = rng.normal(size=(n, n, n))
    density = rng.random(field.shape)
    mesh = mpr.extract_isosurface(field, iso_value=0.1, density=density)
    v, f, d = oracle.extract_isosurface(field, 0.1, density)
    assert len(v) > 0
    assert np.array_equal(mesh.vertices, v)
    assert np.array_equal(mesh.faces, f)
    assert np.array_equal(mesh.density, d)


@pytest.mark.parametrize("n,iso", [(17, 0.0), (34, 0.0), (33, 1.0), (33, -1.0)])
def test_extract_with_iso_ties_bit_exact(n, iso):
    # values in {-1, 0, 1}: many corners equal the iso value, which counts as not below it
    # (capi.mojo:839-1041); roots on edges that end exactly at the iso value
    rng = np.random.default_rng(2000 + n)
    field = rng.integers(-1, 2, size=(n, n, n)).astype(np.float64)
    density = rng.random(field.shape)
    mesh = mpr.extract_isosurface(field, iso_value=iso, density=density)
    v, f, d = oracle.extract_isosurface(field, iso, density)
    assert np.array_equal(mesh.vertices, v)
    assert np.array_equal(mesh.faces, f)
    assert np.array_equal(mesh.density, d)


def test_extract_empty_and_full():
    field = np.ones((9, 9, 9))
    assert mpr.extract_isosurface(field, iso_value=0.0).is_empty
    assert mpr.extract_isosurface(field, iso_value=2.0).is_empty


def test_reconstruct_matches_oracle_on_the_oracle_field():
    p, n = shapes.sphere(1200, 2)
    got = mpr.reconstruct(p, n, depth=5, cycles=5)
    want = oracle.reconstruct(p, n, depth=5, cycles=5)
    assert got.iso_value == pytest.approx(want["iso_value"], rel=1e-10)
    assert got.residual_norm == pytest.approx(want["residual_norm"], rel=1e-8)
    # the solved fields agree to ~1e-12; the meshes then agree vertex for vertex unless a
    # node sits within that distance of the iso value, which these seeds do not produce
    assert got.faces.shape == want["faces"].shape
    assert np.array_equal(got.faces, want["faces"])
    assert np.allclose(got.vertices, want["vertices"], rtol=0, atol=1e-9)
    assert np.allclose(got.density, want["density"], rtol=1e-12, atol=1e-12)


def test_sphere_invariants():
    # reference tests/test_reconstruction.py:65-83
    p, _ = shapes.sphere(1200, 2)
    mesh = mpr.reconstruct(p, p, depth=5, cycles=5)
    edges = np.sort(np.concatenate([mesh.faces[:, [0, 1]], mesh.faces[:, [1, 2]],
                                    mesh.faces[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    ext = mesh.vertices.max(axis=0) - mesh.vertices.min(axis=0)
    assert ext.max() / ext.min() < 1.04
    v, f = mesh.vertices, mesh.faces
    vol = abs(np.einsum("ij,ij->i", v[f[:, 0]], np.cross(v[f[:, 1]], v[f[:, 2]])).sum() / 6)
    assert vol == pytest.approx(4 * np.pi / 3, rel=0.04)
    assert mesh.residual_norm / np.linalg.norm(p) < 0.2


def test_reconstruct_mesh_paths():
    # reference tests/test_reconstruction.py:147-168
    assert mpr.reconstruct_mesh(np.empty((0, 3)), np.empty((0, 3), np.int64)).is_empty
    zero_area = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0]]),
                 np.array([[0, 1, 2]]))
    assert mpr.reconstruct_mesh(*zero_area).is_empty
    vertices = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
                        dtype=np.float64)
    faces = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5],
                      [3, 1, 5], [0, 3, 5]], dtype=np.int64)
    mesh = mpr.reconstruct_mesh(vertices, faces, depth=3, cycles=2)
    assert not mesh.is_empty and mesh.faces.shape[1] == 3


def test_job_reuse_with_smaller_input_and_graph_replay():
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job
    p, n = shapes.sphere(6000, 1)
    unit, _, _ = oracle.normalize(p, 1.1)
    job = Job(6, 6000)
    params = Job.params(_lib.STAGE_SOLVE | _lib.STAGE_EXTRACT)
    outs = []
    for count in (6000, 2500):
        job.set_points(unit[:count], n[:count])
        job.run(params)
        r = job.wait()
        outs.append((r.iso_value, job.fetch_mesh(r)))
        want = oracle.solve(unit[:count], n[:count], 6)
        assert r.iso_value == pytest.approx(want["iso_value"], rel=1e-10)
    # capture the same run as a HIP graph and replay it several times
    job.capture(params)
    for _ in range(3):
        job.launch()
        r = job.wait()
        assert r.iso_value == pytest.approx(outs[1][0], rel=1e-12)
        v, f, _ = job.fetch_mesh(r)
        assert np.array_equal(f, outs[1][1][1])
        assert np.allclose(v, outs[1][1][0], rtol=0, atol=1e-12)
    job.close()


def test_concurrent_calls_from_threads_match_single_calls():
    # the public calls may run in several host threads at once, as the reference's stateless
    # calls can: each call gets a device job of its own, and every result equals the same
    # call made alone
    from concurrent.futures import ThreadPoolExecutor
    clouds = [shapes.sphere(3000, s) for s in range(3)] + [shapes.torus(4000, s) for s in range(3)]

    def call(i):
        p, n = clouds[i % len(clouds)]
        return mpr.reconstruct(p, n, depth=6, precision="mixed" if i % 2 else "fp64")

    with ThreadPoolExecutor(4) as pool:
        got = list(pool.map(call, range(3 * len(clouds))))
    for i, m in enumerate(got):
        want = call(i)
        assert m.iso_value == want.iso_value and m.residual_norm == want.residual_norm, i
        assert np.array_equal(m.vertices, want.vertices), i
        assert np.array_equal(m.faces, want.faces), i
        assert np.array_equal(m.density, want.density), i
    units = [oracle.normalize(p, 1.1)[0] for p, _ in clouds]
    with ThreadPoolExecutor(3) as pool:
        trees = list(pool.map(lambda u: mpr.build_octree(u, 6), units))
    for u, t in zip(units, trees):
        want = mpr.build_octree(u, 6)
        assert np.array_equal(t.morton, want.morton) and np.array_equal(t.parent, want.parent)


def test_job_mesh_overflow_regrows_and_matches_extraction():
    # a noise cloud whose mesh outgrows the job's preallocated vertex buffer: the job grows
    # its buffers and re-extracts; the result equals a standalone extraction of its field
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job

    rng = np.random.default_rng(9)
    depth, count = 6, 30000
    points = rng.random((count, 3)) * 0.9 + 0.05
    normals = rng.normal(size=(count, 3))
    n = (1 << depth) + 1
    vcap = min((n - 1) ** 3, max(1 << 16, 12 * n * n))
    job = Job(depth, count)
    try:
        job.set_points(points, normals)
        job.run(Job.params(_lib.STAGE_SOLVE | _lib.STAGE_EXTRACT | _lib.STAGE_GRIDS))
        r = job.wait()
        assert r.vertex_count > vcap
        v, f, d = job.fetch_mesh(r)
        field, density = job.fetch_grids()[:2]
    finally:
        job.close()
    mesh = mpr.extract_isosurface(field, iso_value=r.iso_value, density=density)
    assert np.array_equal(v, mesh.vertices)
    assert np.array_equal(f, mesh.faces)
    assert np.array_equal(d, mesh.density)


@pytest.mark.parametrize("precision", [0, 1])
def test_sparse_assembly_over_a_sequence_equals_fresh_dense_runs(precision):
    # Without STAGE_GRIDS a job's assembly writes only the nodes whose stencil meets this or
    # the previous run's splat (every other node keeps the +0.0 an earlier run wrote) and the
    # extraction reads the raw splat density through the splat's node bits. A sequence of
    # different clouds on one job, with a dense STAGE_GRIDS run, an empty cloud and a graph
    # replay in between, must give bit for bit what a fresh job's dense run gives.
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job
    depth = 6
    base = _lib.STAGE_SOLVE | _lib.STAGE_EXTRACT
    clouds = []
    for maker, count, seed in ((shapes.sphere, 4000, 1), (shapes.torus, 6000, 2),
                               (shapes.superquadric, 5000, 3)):
        p, n = maker(count, seed=seed)
        clouds.append((oracle.normalize(p, 1.1)[0], n))
    empty = (np.zeros((0, 3)), np.zeros((0, 3)))
    seq = [(clouds[0], 0), (clouds[1], 0), (empty, 0), (clouds[2], _lib.STAGE_GRIDS),
           (clouds[1], 0), (clouds[0], 0)]

    def dense_fresh(cloud, params_grids):
        j = Job(depth, 6000, precision)
        try:
            j.set_points(*cloud)
            j.run(params_grids)
            r = j.wait()
            return r, j.fetch_mesh(r)
        finally:
            j.close()

    job = Job(depth, 6000, precision)
    try:
        for cloud, extra in seq:
            job.set_points(*cloud)
            job.run(Job.params(base | extra, confidence=True))
            r = job.wait()
            v, f, d = job.fetch_mesh(r)
            rw, (vw, fw, dw) = dense_fresh(cloud, Job.params(base | _lib.STAGE_GRIDS,
                                                             confidence=True))
            assert r.iso_value == rw.iso_value and r.residual_norm == rw.residual_norm
            assert np.array_equal(v, vw) and np.array_equal(f, fw) and np.array_equal(d, dw)
            if not len(cloud[0]):
                assert len(v) == 0
        # a sparse graph recorded now, an eager dense run with other samples_per_node, then
        # the replay: the replay's mesh is the sparse run's (its density divided by its own
        # samples_per_node)
        job.set_points(*clouds[2])
        sparse = Job.params(base, samples_per_node=2.0)
        job.run(sparse)
        r0 = job.wait()
        want = job.fetch_mesh(r0)
        job.capture(sparse)
        job.run(Job.params(base | _lib.STAGE_GRIDS, samples_per_node=1.0))
        job.wait()
        for _ in range(2):
            job.launch()
            r = job.wait()
            got = job.fetch_mesh(r)
            assert r.iso_value == r0.iso_value
            assert all(np.array_equal(a, b) for a, b in zip(got, want))
    finally:
        job.close()


def test_mixed_runs_with_and_without_the_initial_fill_match_fresh_jobs():
    # A mixed-precision run whose first cycle fuses the correction into the FP32 sweep
    # (post-smoothing, a grid above the LDS-resident classes) does not fill the FP64
    # solution with the initial guess 0.5: the first defect reads it as 0.5 and the first
    # correction stores 0.5 + e. Runs on one job that take the filled path (no
    # post-smoothing, no cycles) and runs that do not, in any order, give bit for bit what
    # a fresh job gives with the same parameters.
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job
    p, n = shapes.torus(5000, seed=4)
    unit = oracle.normalize(p, 1.1)[0]
    stages = _lib.STAGE_SOLVE | _lib.STAGE_EXTRACT
    runs = [{}, {"post_smooth": 0}, {"post_smooth": 2, "cycles": 1}, {"cycles": 0},
            {"pre_smooth": 0}, {}]

    def run(job, kw):
        job.set_points(unit, n)
        job.run(Job.params(stages, **kw))
        r = job.wait()
        return r, job.fetch_mesh(r)

    job = Job(6, 5000, _lib.PRECISION_MIXED)
    try:
        for kw in runs:
            r, mesh = run(job, kw)
            fresh = Job(6, 5000, _lib.PRECISION_MIXED)
            try:
                rw, want = run(fresh, kw)
            finally:
                fresh.close()
            assert r.iso_value == rw.iso_value and r.residual_norm == rw.residual_norm, kw
            assert all(np.array_equal(a, b) for a, b in zip(mesh, want)), kw
            if kw.get("cycles", 4):
                assert len(mesh[0]) > 0
    finally:
        job.close()


FOOTPRINT_PROBE = """
import ctypes, json, sys
sys.path.insert(0, sys.argv[1])
from mojo_poissonrecon import _lib
from mojo_poissonrecon.core import Job
lib = _lib.lib()
free = ctypes.c_int64()
def free_bytes():
    _lib.check(lib.mpr_device_memory(ctypes.byref(free), None), "mpr_device_memory")
    return free.value
Job(4, 100).close()  # the runtime's one-time allocations come before the measurement
out, jobs = [], []
for depth, count, precision in ((4, 100, 0), (6, 20000, 0), (8, 200000, 1)):
    before = free_bytes()
    jobs.append(Job(depth, count, precision))
    out.append((Job.footprint(depth, count, precision), before - free_bytes()))
print(json.dumps(out))
"""


def test_job_holds_exactly_its_footprint():
    # a job's buffers are carved out of one device allocation of mpr_job_footprint bytes, so
    # admission by footprint (the server) counts what the device gives away, up to the
    # rounding of that one allocation and the job stream's own memory. Measured in a fresh
    # process with every job kept, so that no memory freed earlier is handed out again.
    import json
    import os
    import subprocess
    import sys
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(mpr.__file__)))
    run = subprocess.run([sys.executable, "-c", FOOTPRINT_PROBE, pkg], capture_output=True,
                         text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    # besides the one allocation, a job's HIP stream takes a few MiB of its own
    pairs = json.loads(run.stdout.strip().splitlines()[-1])
    assert all(fp <= used <= fp + (8 << 20) for fp, used in pairs), pairs


def test_fetch_grids_needs_a_grids_run():
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job
    p, n = shapes.sphere(2000, 4)
    job = Job(5, 2000)
    try:
        job.set_points(oracle.normalize(p, 1.1)[0], n)
        job.run(Job.params(_lib.STAGE_SOLVE))
        job.wait()
        with pytest.raises(_lib.DeviceError, match="STAGE_GRIDS"):
            job.fetch_grids()
        job.run(Job.params(_lib.STAGE_SOLVE | _lib.STAGE_GRIDS))
        job.wait()
        field, density, screen, rhs = job.fetch_grids()
        assert density.sum() > 0 and np.isfinite(field).all()
    finally:
        job.close()


def test_job_rejects_negative_or_non_finite_point_weight():
    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job
    job = Job(4, 16)
    try:
        for pw in (-1.0, float("nan"), float("inf")):
            with pytest.raises(_lib.DeviceError):
                job.run(Job.params(_lib.STAGE_SOLVE, point_weight=pw))
    finally:
        job.close()


def test_host_mesh_output_equals_device_fetch_and_round_trips_as_ply(tmp_path):
    import ctypes

    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job, mesh_from_job, prepare_reconstruct
    from mojo_poissonrecon.ply import read_ply, write_ply

    p, n = shapes.torus_workload(50_000, seed=5)
    req = prepare_reconstruct(p, n, depth=7, confidence=True)
    dev = Job(7, len(req.points), req.precision, host_mesh=False)
    host = Job(7, len(req.points), req.precision, host_mesh=True)
    meshes = []
    for job in (dev, host):
        job.set_points(req.points, req.normals)
        job.run(req.params)
        meshes.append(mesh_from_job(job, job.wait(), len(req.points)))
    md, mh = meshes
    assert len(mh.faces) > 10_000
    assert np.array_equal(md.vertices, mh.vertices)
    assert np.array_equal(md.faces, mh.faces)
    assert np.array_equal(md.density, mh.density)
    # the pinned buffers themselves hold the mesh
    v, d, f = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(_lib.lib().mpr_job_host_mesh(host.handle, ctypes.byref(v), ctypes.byref(d),
                                            ctypes.byref(f)), "mpr_job_host_mesh")
    pv = np.ctypeslib.as_array(ctypes.cast(v, ctypes.POINTER(ctypes.c_double)),
                               (len(mh.vertices), 3))
    pf = np.ctypeslib.as_array(ctypes.cast(f, ctypes.POINTER(ctypes.c_int64)), (len(mh.faces), 3))
    assert np.array_equal(pv, mh.vertices) and np.array_equal(pf, mh.faces)
    # graph replay in host mode, then host mode off and on again (graphs re-record)
    host.capture(req.params)
    host.launch()
    assert np.array_equal(mesh_from_job(host, host.wait(), len(req.points)).faces, mh.faces)
    host.set_host_mesh(False)
    host.launch()
    assert np.array_equal(mesh_from_job(host, host.wait(), len(req.points)).vertices, mh.vertices)
    assert _lib.lib().mpr_job_host_mesh(host.handle, None, None, None) == -1
    host.set_host_mesh(True)
    host.run(req.params)
    assert np.array_equal(mesh_from_job(host, host.wait(), len(req.points)).faces, mh.faces)
    write_ply(tmp_path / "torus.ply", mh)
    rv, rf, rd = read_ply(tmp_path / "torus.ply")
    assert np.array_equal(rv, mh.vertices) and np.array_equal(rf, mh.faces)
    assert np.array_equal(rd, mh.density)
    dev.close()
    host.close()


def adversarial_rows(count=20000, seed=0):
    # core.py:73-83 with huge, tiny, zero, infinite and NaN components
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(count, 3)) * 10.0 ** rng.integers(-300, 300, size=(count, 1))
    N = rng.normal(size=(count, 3)) * 10.0 ** rng.integers(-320, 308, size=(count, 1))
    N[::7] = 0
    N[1::11, 0] = np.inf
    P[2::13, 1] = np.nan
    N[3::17] = [0, 0, 5e-324]
    N[4::19] = [1.7e308, -1.7e308, 1e308]
    N[5::23] = [-0.0, 0.0, -0.0]
    P[6::29, 2] = -np.inf
    return P, N


def test_device_row_filter_and_normalisation_equal_the_reference_restatement():
    import ctypes

    from mojo_poissonrecon import _lib
    from mojo_poissonrecon.core import Job

    P, N = adversarial_rows()
    with np.errstate(all="ignore"):
        mask = oracle.valid_oriented_rows(P, N)
        unit, origin, width = oracle.normalize(P[mask], 1.1)
    job = Job(5, len(P), _lib.PRECISION_FP64)
    job.set_points(P, N)
    job.run(Job.params(_lib.STAGE_FILTER | _lib.STAGE_NORMALIZE))
    r = job.wait()
    count = ctypes.c_int64()
    got_unit, got_n = np.zeros_like(P), np.zeros_like(N)
    _lib.check(_lib.lib().mpr_job_fetch_filtered(job.handle, ctypes.byref(count),
                                                 got_unit.ctypes.data, got_n.ctypes.data),
               "mpr_job_fetch_filtered")
    assert count.value == mask.sum() == r.valid_rows
    assert np.array_equal(got_n[:count.value], N[mask])  # kept rows, in order, bit for bit
    assert np.array_equal(got_unit[:count.value], unit)
    assert list(r.origin) == list(origin) and r.width == width
    # a cloud whose every row is dropped: the identity transform and no kept row
    job.set_points(P[~mask][:500], N[~mask][:500])
    job.run(Job.params(_lib.STAGE_FILTER | _lib.STAGE_NORMALIZE))
    r = job.wait()
    assert r.valid_rows == 0 and list(r.origin) == [0.0, 0.0, 0.0] and r.width == 1.0
    job.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reconstruct_with_invalid_rows_equals_reconstruct_of_the_kept_rows(precision):
    from mojo_poissonrecon.batch import reconstruct_batch

    p, n = shapes.torus_workload(40_000, seed=9)
    rng = np.random.default_rng(2)
    bad_p = rng.uniform(-1, 1, (6000, 3))
    bad_n = rng.normal(size=(6000, 3))
    bad_p[0::6, 1] = np.nan
    bad_n[1::6, 0] = np.inf
    bad_n[2::6] = 0.0
    bad_n[3::6] = [-0.0, 0.0, -0.0]
    bad_p[4::6, 2] = -np.inf
    bad_n[5::6] = [0.0, 0.0, 5e-324]  # kept: a subnormal but non-zero normal
    order = rng.permutation(len(p) + len(bad_p))
    allp = np.concatenate([p, bad_p])[order]
    alln = np.concatenate([n, bad_n])[order]
    with np.errstate(all="ignore"):
        keep = oracle.valid_oriented_rows(allp, alln)
    assert keep.sum() == len(p) + 1000
    want = mpr.reconstruct(np.ascontiguousarray(allp[keep]), np.ascontiguousarray(alln[keep]),
                           depth=7, confidence=True, precision=precision)
    got = mpr.reconstruct(allp, alln, depth=7, confidence=True, precision=precision)
    for a, b in ((got, want),):
        assert a.iso_value == b.iso_value and a.residual_norm == b.residual_norm
        assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.faces, b.faces)
        assert np.array_equal(a.density, b.density)
    # the same inside a batch, beside a cloud with no valid row
    none_p, none_n = rng.uniform(-1, 1, (300, 3)), np.zeros((300, 3))
    batched = reconstruct_batch([(allp, alln), (none_p, none_n)], depths=[7, 6], confidence=True,
                                precision=precision)
    assert np.array_equal(batched[0].faces, want.faces)
    assert np.array_equal(batched[0].vertices, want.vertices)
    assert batched[1].is_empty and batched[1].iso_value == 0.0
    # no valid row: an empty mesh, before the solver parameters are looked at (core.py:410-411)
    assert mpr.reconstruct(none_p, none_n, depth=6).is_empty
    assert mpr.reconstruct(none_p, none_n, depth=99).is_empty
    with pytest.raises(ValueError, match="depth"):
        mpr.reconstruct(p, n, depth=99)
