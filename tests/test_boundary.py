"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every symbol its
header declares, the buffer conversion rules and validation messages match the reference,
and the host-side mesh preprocessing matches the reference's fixtures. No GPU calls."""

from __future__ import annotations

import os
import re

import numpy as np
import pytest

import mojo_poissonrecon as mpr
import oracle
from mojo_poissonrecon import _lib
from mojo_poissonrecon._lib import addr, f64, i64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mpr_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    handle = _lib.load_library()
    names = header_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(handle, name), name
    # the reference's kernel table is a subset (python/mojo_poissonrecon/_lib.py:20-32)
    for name in _lib._SIGNATURES:
        assert name in names


def test_reference_kernel_signatures_unchanged():
    I, F = _lib.I, _lib.F
    assert _lib._SIGNATURES["mpr_splat_f64"] == ([I] * 10 + [F], I)
    assert _lib._SIGNATURES["mpr_residual_f64"] == ([I, I, I, I, I, F], F)
    assert _lib._SIGNATURES["mpr_dual_vertices_f64"] == ([I, I, I, F, I, I, I, I], I)


def test_job_footprint_is_computed_without_a_device():
    fp8 = mpr.Job.footprint(8, 200_000)
    fp8m = mpr.Job.footprint(8, 200_000, _lib.PRECISION_MIXED)
    n = 257
    assert fp8 > 7 * 8 * n ** 3  # at least the seven natural-layout float64 grids
    assert fp8m > fp8
    assert mpr.Job.footprint(11, 10) == -1


def test_ffi_conversions_reject_narrowing_and_bad_buffers():
    # reference tests/test_kernels.py:91-103
    with pytest.raises(TypeError, match="complex"):
        f64(np.array([1 + 2j]))
    with pytest.raises(TypeError, match="real numbers"):
        f64(["1.0"])
    with pytest.raises(ValueError, match="exactly"):
        f64(np.array([2**53 + 1], dtype=np.int64))
    with pytest.raises(ValueError, match="finite integers"):
        i64(np.array([1.5]))
    with pytest.raises(TypeError, match="dtype"):
        addr(np.ones(3, dtype=np.float32))
    with pytest.raises(ValueError, match="C-contiguous"):
        addr(np.ones((3, 3), dtype=np.float64)[:, ::2])


def test_public_validation_rejects_unsafe_values():
    # reference tests/test_kernels.py:198-208 (raised before any device work)
    points = np.array([[0.25, 0.25, 0.25]])
    normals = np.array([[1.0, 0.0, 0.0]])
    with pytest.raises(TypeError, match="depth"):
        mpr.solve_screened_poisson(points, normals, depth=2.5)
    with pytest.raises(ValueError, match="finite"):
        mpr.solve_screened_poisson(points, normals, point_weight=np.nan)
    with pytest.raises(ValueError, match="density must be finite"):
        mpr.extract_isosurface(np.zeros((2, 2, 2)), density=np.full((2, 2, 2), np.nan))
    with pytest.raises(ValueError, match="greater than 1"):
        mpr.reconstruct(points, normals, scale=1.0)
    with pytest.raises(ValueError, match="depth must be between"):
        mpr.solve_screened_poisson(points, normals, depth=11)
    with pytest.raises(ValueError, match="precision"):
        mpr.solve_screened_poisson(points, normals, precision="fp16")
    with pytest.raises(ValueError, match="inner_cycles"):
        mpr.solve_screened_poisson(points, normals, precision="mixed", inner_cycles=-1)
    with pytest.raises(TypeError, match="inner_cycles"):
        mpr.reconstruct(points, normals, precision="mixed", inner_cycles=True)
    with pytest.raises(ValueError, match="inside"):
        mpr.solve_screened_poisson(points + 1.0, normals)
    with pytest.raises(ValueError, match="cubic"):
        mpr.extract_isosurface(np.zeros((2, 3, 2)))
    with pytest.raises(ValueError, match="octree points"):
        mpr.build_octree(points - 1.0)


EMPTY = (np.empty((0, 3)), np.empty((0, 3), dtype=np.int64), 0)
SINGLE_TRIANGLE = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]),
                   np.array([[0, 1, 2]]), 3)
DUPLICATE_VERTICES = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                                [0.0, 0.0, 0.0]]), np.array([[0, 1, 2], [3, 1, 2]]), 4)
ZERO_AREA = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0]]),
             np.array([[0, 1, 2]]), 0)
NON_MANIFOLD = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.5, -1.0, 0.0],
                          [0.5, 0.0, 1.0]]), np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]]), 5)
UNREFERENCED = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [9.0, 9.0, 9.0]]),
                np.array([[0, 1, 2]]), 3)


@pytest.mark.parametrize("vertices,faces,expected",
                         [EMPTY, SINGLE_TRIANGLE, DUPLICATE_VERTICES, ZERO_AREA, NON_MANIFOLD,
                          UNREFERENCED])
def test_oracle_mesh_preprocessing_on_the_reference_fixtures(vertices, faces, expected):
    # reference tests/test_reconstruction.py:86-144: the oracle's per-face loop, in Python and
    # in C, keeps the expected vertices with unit normals (the GPU form is checked against
    # both in tests/test_gpu_mesh_normals.py)
    points, normals = oracle.oriented_points_from_mesh(vertices, faces)
    assert len(points) == expected
    assert points.shape == normals.shape
    assert np.isfinite(normals).all()
    if expected:
        assert np.allclose(np.linalg.norm(normals, axis=1), 1.0)
    cp, cn = oracle.oriented_points_from_mesh_c(vertices, faces)
    assert np.array_equal(points, cp) and np.array_equal(normals, cn)


def test_oracle_mesh_normals_are_area_weighted():
    # reference tests/test_reconstruction.py:171-179
    vertices = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [0, 0, 1]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [0, 3, 1]], dtype=np.int64)
    for points, normals in (oracle.oriented_points_from_mesh(vertices, faces),
                            oracle.oriented_points_from_mesh_c(vertices, faces)):
        origin = np.flatnonzero(np.all(points == vertices[0], axis=1))[0]
        assert np.allclose(normals[origin], [0.0, 1 / np.sqrt(5), 2 / np.sqrt(5)])


def test_c_mesh_normals_loop_equals_the_python_loop_bitwise():
    rng = np.random.default_rng(5)
    vertices = rng.normal(size=(3000, 3)) * 10.0 ** rng.integers(-160, 160, size=(3000, 1))
    faces = rng.integers(0, 3000, size=(10000, 3))
    faces[:10, 1] = faces[:10, 0]  # degenerate faces
    with np.errstate(all="ignore"):
        p, n = oracle.oriented_points_from_mesh_c(vertices, faces)
        wp, wn = oracle.oriented_points_from_mesh(vertices, faces)
    assert np.array_equal(p, wp) and np.array_equal(n, wn)


def test_mesh_validation_messages():
    v = np.zeros((3, 3))
    with pytest.raises(ValueError, match="shape"):
        mpr.oriented_points_from_mesh(v, np.zeros((2, 4), np.int64))
    with pytest.raises(ValueError, match="out of range"):
        mpr.oriented_points_from_mesh(v, np.array([[0, 1, 3]]))
    with pytest.raises(ValueError, match="finite integers"):
        mpr.oriented_points_from_mesh(v, np.array([[0.5, 1, 2]]))


def test_row_filter_matches_reference_restatement_on_adversarial_rows():
    # core.py:73-83 with huge, tiny, zero, infinite and NaN components
    rng = np.random.default_rng(0)
    P = rng.normal(size=(20000, 3)) * 10.0 ** rng.integers(-300, 300, size=(20000, 1))
    N = rng.normal(size=(20000, 3)) * 10.0 ** rng.integers(-320, 308, size=(20000, 1))
    N[::7] = 0
    N[1::11, 0] = np.inf
    P[2::13, 1] = np.nan
    N[3::17] = [0, 0, 5e-324]
    N[4::19] = [1.7e308, -1.7e308, 1e308]
    N[5::23] = [-0.0, 0.0, -0.0]
    with np.errstate(all="ignore"):
        got = mpr.core._valid_oriented_rows(P, N)
        want = oracle.valid_oriented_rows(P, N)
    assert np.array_equal(got, want)
    clean = rng.normal(size=(100, 3))
    assert mpr.core._valid_oriented_rows(clean, clean).all()


def test_server_job_capacity_rounds_to_powers_of_two():
    from mojo_poissonrecon.serve import job_capacity
    assert job_capacity(0) == 16384 and job_capacity(3000) == 16384
    assert job_capacity(16384) == 16384 and job_capacity(16385) == 32768
    assert job_capacity(500_000) == 524288


def test_workgroup_size_table_without_a_device():
    from mojo_poissonrecon import tune
    from mojo_poissonrecon._lib import DeviceError
    assert tune.tuned_block("relax", "fp32", 8) == 256
    tune.set_tuned_block("relax", "fp32", 8, 512)
    try:
        assert tune.tuned_block("relax", "fp32", 8) == 512
        assert tune.tuned_block("relax", "fp64", 8) == 256
    finally:
        tune.set_tuned_block("relax", "fp32", 8, 256)
    for bad in (0, 32, 200, 2048):
        with pytest.raises(DeviceError):
            tune.set_tuned_block("residual", "fp64", 6, bad)
    with pytest.raises(DeviceError):
        tune.tuned_block("defect", "fp64", 11)
    assert tune.tuned_groups("relax", "fp32", 8) == 1
    tune.set_tuned_groups("relax", "fp32", 8, 2)
    try:
        assert tune.tuned_groups("relax", "fp32", 8) == 2
    finally:
        tune.set_tuned_groups("relax", "fp32", 8, 1)
    for kernel, groups in (("relax", 3), ("relax", 0), ("residual", 2)):
        with pytest.raises(DeviceError):
            tune.set_tuned_groups(kernel, "fp32", 8, groups)
    assert tune.tuned_planes("relax", "fp64", 7) == 1
    tune.set_tuned_planes("relax", "fp64", 7, 4)
    try:
        assert tune.tuned_planes("relax", "fp64", 7) == 4
        assert tune.tuned_planes("relax", "fp32", 7) == 1
    finally:
        tune.set_tuned_planes("relax", "fp64", 7, 1)
    for kernel, planes in (("relax", 3), ("relax", 8), ("relax", 0), ("defect", 2)):
        with pytest.raises(DeviceError):
            tune.set_tuned_planes(kernel, "fp64", 7, planes)


def test_ply_round_trip(tmp_path):
    from mojo_poissonrecon.core import Mesh
    from mojo_poissonrecon.ply import read_ply, write_ply

    rng = np.random.default_rng(4)
    v = rng.normal(size=(50, 3)) * 1e3
    f = rng.integers(0, 50, (80, 3))
    d = rng.random(50)
    mesh = Mesh(v, f, d, 0.5, 1.0)
    for binary in (True, False):
        path = tmp_path / f"m{int(binary)}.ply"
        write_ply(path, mesh, binary=binary)
        rv, rf, rd = read_ply(path)
        assert np.array_equal(rv, v) and np.array_equal(rf, f) and np.array_equal(rd, d)
        header = path.read_bytes().split(b"end_header")[0].decode()
        assert "element vertex 50" in header and "element face 80" in header
        assert "property list uchar int vertex_indices" in header
    write_ply(tmp_path / "nd.ply", mesh, density=False)
    rv, rf, rd = read_ply(tmp_path / "nd.ply")
    assert rd is None and np.array_equal(rv, v)
    # binary layout: 4 doubles per vertex, then 13 bytes per face
    size = (tmp_path / "m1.ply").stat().st_size
    head = len((tmp_path / "m1.ply").read_bytes().split(b"end_header\n")[0]) + len(b"end_header\n")
    assert size == head + 50 * 32 + 80 * 13
    empty = Mesh(np.empty((0, 3)), np.empty((0, 3), np.int64), np.empty(0))
    write_ply(tmp_path / "e.ply", empty)
    rv, rf, _ = read_ply(tmp_path / "e.ply")
    assert rv.shape == (0, 3) and rf.shape == (0, 3)
    with pytest.raises(ValueError, match="out of range"):
        write_ply(tmp_path / "bad.ply", Mesh(v, f + 50, d))
