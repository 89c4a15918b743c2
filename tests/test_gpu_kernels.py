"""GPU parity of the drop-in kernel ABI (the reference's eleven symbols) against the oracle.

Each test calls libmpr_hip.so through ctypes exactly as the reference's own tests call its
library (tests/test_kernels.py of the reference), on seeded inputs, and compares with the
CPU oracle: bit for bit where the GPU evaluates the same operations in the same order
(relaxation, residual values, restriction, prolongation, assembly, Morton codes, dual
vertices and faces), and within stated tolerances where the GPU sums in another order
(splat atomics, the residual norm, the iso-value reduction).
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from mojo_poissonrecon._lib import addr, lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rng_grids(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n, n)), rng.normal(size=(n, n, n)), rng.random((n, n, n))


@pytest.mark.parametrize("n", [2, 9, 10, 33, 34, 65, 66, 129])
def test_relax_bit_exact(n):
    x, rhs, screen = rng_grids(n, 71)
    expected = oracle.relax(x, rhs, screen, 4.0, 3)
    actual = x.copy()
    lib().mpr_relax_f64(addr(actual), addr(rhs), addr(screen), n, 4.0, 3)
    assert np.array_equal(actual, expected)


@pytest.mark.parametrize("n", [9, 34, 65, 129])
def test_residual_bit_exact_and_norm(n):
    x, rhs, screen = rng_grids(n, n)
    expected, norm = oracle.residual(x, rhs, screen, 4.0)
    actual = np.empty_like(x)
    got = lib().mpr_residual_f64(addr(x), addr(rhs), addr(screen), addr(actual), n, 4.0)
    assert np.array_equal(actual, expected)
    assert got == pytest.approx(norm, rel=1e-14)
    assert got == pytest.approx(np.linalg.norm(expected), rel=2e-15)


def f32_grids(n, seed):
    x, rhs, screen = rng_grids(n, seed)
    return x.astype(np.float32), rhs.astype(np.float32), screen.astype(np.float32)


@pytest.mark.parametrize("n", [3, 9, 17, 33, 34, 65, 66])
def test_relax_f32_bit_exact_against_float32_restatement(n):
    # the mixed-precision smoother: scalar kernels below n = 33, 16-byte slot groups above
    x, rhs, screen = f32_grids(n, 300 + n)
    expected = oracle.relax_np(x, rhs, screen, 4.0, 3)
    actual = x.copy()
    assert lib().mpr_relax_f32(actual.ctypes.data, rhs.ctypes.data, screen.ctypes.data, n, 4.0,
                               3) == 0
    assert np.array_equal(actual, expected)


@pytest.mark.parametrize("n", [9, 33, 34, 65])
def test_residual_f32_bit_exact_against_float32_restatement(n):
    x, rhs, screen = f32_grids(n, 400 + n)
    expected = oracle.residual_np(x, rhs, screen, 4.0)
    actual = np.empty_like(x)
    assert lib().mpr_residual_f32(x.ctypes.data, rhs.ctypes.data, screen.ctypes.data,
                                  actual.ctypes.data, n, 4.0) == 0
    assert np.array_equal(actual, expected)


@pytest.mark.parametrize("n", [3, 9, 33, 65, 66, 129])
def test_restrict_f32_bit_exact_against_float32_restatement(n):
    fine = np.random.default_rng(500 + n).normal(size=(n, n, n)).astype(np.float32)
    cn = (n + 1) // 2
    coarse = np.empty((cn, cn, cn), dtype=np.float32)
    assert lib().mpr_restrict_f32(fine.ctypes.data, coarse.ctypes.data, n) == 0
    assert np.array_equal(coarse, oracle.restrict_np(fine))


@pytest.mark.parametrize("n", [3, 9, 33, 34, 65, 66, 129])
def test_prolong_f32_bit_exact_against_float32_restatement(n):
    rng = np.random.default_rng(600 + n)
    cn = (n + 1) // 2
    coarse = rng.normal(size=(cn, cn, cn)).astype(np.float32)
    fine = rng.normal(size=(n, n, n)).astype(np.float32)
    expected = oracle.prolong_np(coarse, fine)
    assert lib().mpr_prolong_add_f32(coarse.ctypes.data, fine.ctypes.data, n) == 0
    assert np.array_equal(fine, expected)


@pytest.mark.parametrize("n", [65, 129, 257])
def test_colour1_prolongation_bit_exact(n):
    # the mixed V-cycle's prolongation before a post-smoothing (a lane updates four fine
    # rows from one load of their shared coarse rows): colour 1 gets the float32
    # restatement's sum, colour 0 keeps its values
    rng = np.random.default_rng(900 + n)
    cn = (n + 1) // 2
    coarse = rng.normal(size=(cn, cn, cn)).astype(np.float32)
    coarse[0, 0, 0] = -0.0  # signed zeros and exact halves take the same roundings
    coarse[cn - 1, cn - 1, cn - 1] = 0.5
    fine = rng.normal(size=(n, n, n)).astype(np.float32)
    z, y, x = np.indices(fine.shape)
    want = np.where(((x + y + z) & 1) == 1, oracle.prolong_np(coarse, fine), fine)
    got = fine.copy()
    assert lib().mpr_prolong_c1_f32(coarse.ctypes.data, got.ctypes.data, n) == 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("n", [3, 9, 10, 17, 65, 66, 129, 130])
def test_restrict_bit_exact(n):
    fine = np.random.default_rng(11).normal(size=(n, n, n))
    cn = (n + 1) // 2
    coarse = np.empty((cn, cn, cn))
    lib().mpr_restrict_f64(addr(fine), addr(coarse), n)
    assert np.array_equal(coarse, oracle.restrict(fine))


@pytest.mark.parametrize("n", [3, 9, 10, 33, 34, 65, 66, 129])
def test_prolong_add_bit_exact(n):
    rng = np.random.default_rng(5)
    cn = (n + 1) // 2
    coarse = rng.normal(size=(cn, cn, cn))
    fine = rng.normal(size=(n, n, n))
    expected = oracle.prolong_add(coarse, fine)
    lib().mpr_prolong_add_f64(addr(coarse), addr(fine), n)
    assert np.array_equal(fine, expected)


def test_prolong_known_answer():
    # reference tests/test_kernels.py:67-72
    coarse = np.arange(125, dtype=np.float64).reshape(5, 5, 5)
    fine = np.zeros((9, 9, 9))
    lib().mpr_prolong_add_f64(addr(coarse), addr(fine), 9)
    assert np.array_equal(fine[::2, ::2, ::2], coarse)
    assert fine[1, 1, 1] == pytest.approx(coarse[:2, :2, :2].mean())


@pytest.mark.parametrize("n", [5, 33])
def test_assemble_bit_exact(n):
    rng = np.random.default_rng(n)
    vx, vy, vz = (rng.normal(size=(n, n, n)) for _ in range(3))
    screen = rng.random((n, n, n))
    rhs = np.empty_like(vx)
    lib().mpr_assemble_f64(addr(vx), addr(vy), addr(vz), addr(screen), addr(rhs), n, 4.0, 0.5)
    assert np.array_equal(rhs, oracle.assemble(vx, vy, vz, screen, 4.0, 0.5))


@pytest.mark.parametrize("depth", [1, 7, 9, 19])
def test_octree_codes_bit_exact(depth):
    points = np.random.default_rng(3).random((5000, 3))
    points[:3] = [[0, 0, 0], [1, 1, 1], [0.5, 0.25, 1.0]]
    codes = np.empty((len(points), depth + 1), np.int64)
    lib().mpr_octree_codes_f64(addr(points), len(points), depth, addr(codes))
    assert np.array_equal(codes, oracle.octree_codes(points, depth))


@pytest.mark.parametrize("confidence", [0, 1])
def test_splat_bit_exact(confidence):
    rng = np.random.default_rng(17)
    n = 17
    points = rng.random((4000, 3))
    points[:2] = [[1.0, 1.0, 1.0], [0.0, 0.5, 1.0]]
    normals = rng.normal(size=(4000, 3))
    normals[5] = 0.0  # rejected sample
    grids = [np.zeros((n, n, n)) for _ in range(5)]
    acc = lib().mpr_splat_f64(addr(points), addr(normals), len(points), n, confidence,
                              *(addr(g) for g in grids), 1.5)
    eacc, *expected = oracle.splat(points, normals, n, bool(confidence), 1.5)
    assert acc == eacc == len(points) - 1
    # contributions are summed per node in the reference's point order
    for got, want in zip(grids, expected):
        assert np.array_equal(got, want)


def test_splat_accumulates_into_nonzero_grids_in_reference_order():
    # the reference adds to whatever the caller's grids hold (capi.mojo:102-224)
    rng = np.random.default_rng(18)
    n = 9
    points = rng.random((600, 3))
    normals = rng.normal(size=(600, 3))
    start = [rng.normal(size=(n, n, n)) for _ in range(5)]
    grids = [g.copy() for g in start]
    lib().mpr_splat_f64(addr(points), addr(normals), len(points), n, 1,
                        *(addr(g) for g in grids), 1.5)
    want = [g.copy() for g in start]
    lib_o = oracle.lib()
    lib_o.oracle_splat(oracle._p(points), oracle._p(normals), len(points), n, 1,
                       *(oracle._p(g) for g in want), 1.5)
    for got, exp in zip(grids, want):
        assert np.array_equal(got, exp)


def test_splat_golden_vectors_and_confidence_kat():
    with np.load(os.path.join(GOLDEN, "splat3_depth2.npz"), allow_pickle=False) as g:
        g = dict(g)
    grids = [np.zeros((5, 5, 5)) for _ in range(5)]
    acc = lib().mpr_splat_f64(addr(g["points"]), addr(g["normals"]), 3, 5, 1,
                              *(addr(x) for x in grids), 1.0)
    assert acc == 3
    density, vx, vy, vz, screen = grids
    assert np.allclose(density, g["density"], rtol=0, atol=2e-15)
    assert np.allclose(screen, g["screen"], rtol=0, atol=2e-15)
    for a, b in ((vx, g["vx"]), (vy, g["vy"]), (vz, g["vz"])):
        assert np.allclose(a, b, rtol=0, atol=2e-13)


def test_iso_value_known_answers_and_oracle():
    coarse = np.arange(125, dtype=np.float64).reshape(5, 5, 5)
    fine = oracle.prolong_add(coarse, np.zeros((9, 9, 9)))
    points = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [1.0, 1.0, 1.0]])
    normals = np.ones_like(points)
    iso = lib().mpr_iso_value_f64(addr(fine), addr(points), addr(normals), 3, 9, 0)
    assert iso == pytest.approx(np.mean([fine[0, 0, 0], fine[4, 4, 4], fine[8, 8, 8]]))
    rng = np.random.default_rng(2)
    field = rng.normal(size=(33, 33, 33))
    pts = rng.random((20000, 3))
    nrm = rng.normal(size=(20000, 3))
    for conf in (0, 1):
        got = lib().mpr_iso_value_f64(addr(field), addr(pts), addr(nrm), len(pts), 33, conf)
        assert got == pytest.approx(oracle.iso_value(field, pts, nrm, bool(conf)), rel=1e-12,
                                    abs=1e-15)


@pytest.mark.parametrize("n", [14, 65])
def test_dual_vertex_count_mode(n):
    field = np.random.default_rng(n).normal(size=(n, n, n))
    chunks = ((n - 1) + 3) // 4
    scratch = np.full(chunks + 2, -7, dtype=np.int64)
    dummy = np.zeros(1)
    got = lib().mpr_dual_vertices_f64(addr(field), addr(field), n, 0.0, addr(scratch),
                                      addr(dummy), addr(dummy), 0)
    assert got == oracle.dual_vertex_count(field, 0.0)
    # ABI: exactly ceil(cells/4) partial counts written, nothing beyond
    assert scratch[:chunks].sum() == got and np.all(scratch[chunks:] == -7)


def test_dual_contour_golden_and_oracle_bit_exact():
    with np.load(os.path.join(GOLDEN, "dual_sphere13.npz"), allow_pickle=False) as g:
        g = dict(g)
    field = g["field"]
    n = field.shape[0]
    count = oracle.dual_vertex_count(field, 0.0)
    cm = np.empty((n - 1) ** 3, np.int64)
    v = np.empty((count, 3))
    d = np.empty(count)
    got = lib().mpr_dual_vertices_f64(addr(field), addr(field), n, 0.0, addr(cm), addr(v),
                                      addr(d), count)
    ocount, ocm, ov, od = oracle.dual_vertices(field, field, 0.0, count)
    assert got == ocount == count
    assert np.array_equal(cm, ocm)
    assert np.array_equal(v, ov) and np.array_equal(d, od)
    assert np.allclose(v, g["vertices"], rtol=0, atol=2e-16)
    dummy = np.zeros((1, 3), np.int64)
    nf = lib().mpr_dual_faces_f64(addr(field), n, 0.0, addr(cm), addr(dummy), 0)
    faces = np.empty((nf, 3), np.int64)
    assert lib().mpr_dual_faces_f64(addr(field), n, 0.0, addr(cm), addr(faces), nf) == nf
    assert np.array_equal(faces, g["faces"])


def test_dual_capacity_semantics_match_oracle():
    n = 21
    field = np.random.default_rng(4).normal(size=(n, n, n))
    total = oracle.dual_vertex_count(field, 0.1)
    cap = total // 2
    cm = np.empty((n - 1) ** 3, np.int64)
    v = np.empty((cap, 3))
    d = np.empty(cap)
    got = lib().mpr_dual_vertices_f64(addr(field), addr(field), n, 0.1, addr(cm), addr(v),
                                      addr(d), cap)
    ocount, ocm, ov, od = oracle.dual_vertices(field, field, 0.1, cap)
    assert got == ocount == total
    assert np.array_equal(cm, ocm) and np.array_equal(v, ov) and np.array_equal(d, od)
    fc, _ = oracle.dual_faces(field, 0.1, ocm, 0)
    for capf in (fc, fc - 3, 7):
        faces = np.zeros((max(capf, 1), 3), np.int64)
        got_f = lib().mpr_dual_faces_f64(addr(field), n, 0.1, addr(ocm), addr(faces), capf)
        ofc, of = oracle.dual_faces(field, 0.1, ocm, capf)
        assert got_f == ofc
        written = 2 * min(ofc // 2, capf // 2)
        assert np.array_equal(faces[:written], of[:written])


@pytest.mark.parametrize("n", [65, 129])
def test_colour0_residual_and_colour1_zero_restriction_bit_exact(n):
    # the mixed V-cycle's batched step after pre-smoothing: residual of colour 0 only
    # (colour-1 words stay zero), then the full-weighting restriction with the fine colour 1
    # read as zero; against the float32 restatements
    x, rhs, screen = f32_grids(n, 700 + n)
    res = np.full_like(x, np.nan)
    coarse = np.full(((n + 1) // 2,) * 3, np.nan, dtype=np.float32)
    assert lib().mpr_vcycle_transfer_f32(x.ctypes.data, rhs.ctypes.data, screen.ctypes.data,
                                         res.ctypes.data, coarse.ctypes.data, n, 4.0) == 0
    z, y, xx = np.indices(x.shape)
    colour0 = ((xx + y + z) & 1) == 0
    want = np.where(colour0, oracle.residual_np(x, rhs, screen, 4.0), np.float32(0))
    assert np.array_equal(res, want)
    assert np.array_equal(coarse, oracle.restrict_np(want))
