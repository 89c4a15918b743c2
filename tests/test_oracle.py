"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each test checks one oracle function against something that does not share its code:
the reference tests' own known answers and closed forms (tests/test_kernels.py,
tests/test_reconstruction.py in the reference), golden vectors written from the
reference's NumPy translations (tests/golden/), or mathematics that fixes the answer
(bit interleaving, separable full weighting, exact reproduction of linear functions,
brute force on tiny grids). CPU only.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as data:
        return {k: data[k] for k in data.files}


def stencil_sum_degree(values):
    """Vectorised Neumann 7-point neighbour sum and degree (reference test_kernels.py:106-121)."""
    s = np.zeros_like(values)
    d = np.zeros(values.shape, dtype=np.int64)
    s[:, :, 1:] += values[:, :, :-1]; d[:, :, 1:] += 1
    s[:, :, :-1] += values[:, :, 1:]; d[:, :, :-1] += 1
    s[:, 1:, :] += values[:, :-1, :]; d[:, 1:, :] += 1
    s[:, :-1, :] += values[:, 1:, :]; d[:, :-1, :] += 1
    s[1:, :, :] += values[:-1, :, :]; d[1:, :, :] += 1
    s[:-1, :, :] += values[1:, :, :]; d[:-1, :, :] += 1
    return s, d


def interleave_codes(points, depth):
    """Morton code by bit interleaving of clamp(trunc(p * 2^D)): x bit 0, y bit 1, z bit 2."""
    side = 1 << depth
    q = np.clip(np.trunc(points * side).astype(np.int64), 0, side - 1)
    code = np.zeros(len(points), dtype=np.int64)
    for bit in range(depth):
        for axis in range(3):
            code |= ((q[:, axis] >> bit) & 1) << (3 * bit + axis)
    return np.stack([code >> (3 * (depth - d)) for d in range(depth + 1)], axis=1)


def test_morton_codes_equal_bit_interleave():
    rng = np.random.default_rng(3)
    points = rng.random((200, 3))
    points[:4] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 0.999999]]
    for depth in (1, 5, 7, 9):
        assert np.array_equal(oracle.octree_codes(points, depth), interleave_codes(points, depth))


def test_octree_levels_parent_and_leaf_relations():
    # reference tests/test_kernels.py:10-21 (seed 3, 200 points, depth 7)
    rng = np.random.default_rng(3)
    points = rng.random((200, 3))
    depth = 7
    expected = interleave_codes(points, depth)
    tree = oracle.build_octree(points, depth)
    for d in range(depth + 1):
        assert np.array_equal(tree["morton"][tree["depth"] == d], np.unique(expected[:, d]))
    assert np.array_equal(tree["morton"][tree["point_leaf"]], expected[:, -1])
    inner = tree["depth"] > 0
    assert np.array_equal(tree["morton"][tree["parent"][inner]], tree["morton"][inner] >> 3)
    # child ranges are contiguous and point back at their parent
    for node in np.flatnonzero(tree["child_count"]):
        s, c = tree["child_start"][node], tree["child_count"][node]
        assert np.all(tree["parent"][s:s + c] == node)
    assert tree["child_count"].sum() == len(tree["morton"]) - 1


def test_splat_and_assemble_match_reference_translation_vectors():
    g = _golden("splat3_depth2.npz")
    acc, density, vx, vy, vz, screen = oracle.splat(g["points"], g["normals"], int(g["n"]),
                                                    bool(g["confidence"]),
                                                    float(g["samples_per_node"]))
    assert acc == 3
    density = density / float(g["samples_per_node"])
    # tolerances of the reference test tests/test_kernels.py:41-43
    assert np.allclose(density, g["density"], rtol=0, atol=2e-15)
    assert np.allclose(screen, g["screen"], rtol=0, atol=2e-15)
    assert np.allclose(vx, g["vx"], rtol=0, atol=2e-13)
    assert np.allclose(vy, g["vy"], rtol=0, atol=2e-13)
    assert np.allclose(vz, g["vz"], rtol=0, atol=2e-13)
    rhs = oracle.assemble(vx, vy, vz, screen, float(g["point_weight"]))
    assert np.allclose(rhs, g["rhs"], rtol=0, atol=2e-13)
    sol = oracle.solve(g["points"], g["normals"], 2, confidence=True, cycles=1,
                       samples_per_node=1.0)
    assert np.allclose(sol["rhs"], g["rhs"], rtol=0, atol=2e-13)


def test_confidence_weight_known_answers():
    # reference tests/test_kernels.py:46-58: exact sums
    points = np.array([[0.25, 0.25, 0.25], [0.75, 0.75, 0.75]])
    normals = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    plain = oracle.solve(points, normals, 2, cycles=1, confidence=False, samples_per_node=1.0)
    weighted = oracle.solve(points, normals, 2, cycles=1, confidence=True, samples_per_node=1.0)
    assert plain["density"].sum() == 4.0
    assert weighted["density"].sum() == 6.0
    assert plain["screen"].sum() == 2.0
    assert weighted["screen"].sum() == 3.0


def test_single_point_vector_field_closed_form():
    # one sample: local density = sum_c 2 w_c^2 (w_c trilinear), factor = max(1, local/spn),
    # v_c = -n_hat * w_c * res / factor
    p = np.array([[0.3, 0.55, 0.8]])
    nrm = np.array([[0.0, 3.0, 4.0]])
    n, spn = 9, 0.25
    acc, density, vx, vy, vz, screen = oracle.splat(p, nrm, n, True, spn)
    res = n - 1
    g = p[0] * res
    i = np.floor(g).astype(int)
    f = g - i
    w = {}
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                w[(dz, dy, dx)] = ((f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1])
                                   * (f[2] if dz else 1 - f[2]))
    length = 5.0
    local = sum(2 * length * wc * wc for wc in w.values())
    factor = max(1.0, local / spn)
    for (dz, dy, dx), wc in w.items():
        idx = (i[2] + dz, i[1] + dy, i[0] + dx)
        assert density[idx] == pytest.approx(2 * wc * length, rel=1e-14)
        assert screen[idx] == pytest.approx(wc * length, rel=1e-14)
        vw = wc * length * res / factor
        assert vx[idx] == 0.0
        assert vy[idx] == pytest.approx(-0.6 * vw, rel=1e-14)
        assert vz[idx] == pytest.approx(-0.8 * vw, rel=1e-14)
    assert np.count_nonzero(screen) == 8


def test_assemble_divergence_of_linear_field_is_constant():
    # v = (a x, b y, c z) on the unit cube: central and one-sided differences are exact,
    # so -div v = -(a + b + c) everywhere; screening adds alpha * s * h^-2 * target.
    n = 9
    ax = np.linspace(0, 1, n)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    vx, vy, vz = 2.0 * x, -3.0 * y, 0.5 * z
    screen = np.zeros((n, n, n))
    screen[4, 4, 4] = 1.0
    rhs = oracle.assemble(vx, vy, vz, screen, 4.0, 0.5)
    expected = np.full((n, n, n), -(2.0 - 3.0 + 0.5))
    expected[4, 4, 4] += 4.0 * 1.0 * (n - 1) ** 2 * 0.5
    assert np.allclose(rhs, expected, rtol=0, atol=1e-12)


def test_restriction_is_renormalised_separable_full_weighting():
    # reference tests/test_kernels.py:61-64 case (seed 11, 9^3) and an odd/even mix
    for n, seed in ((9, 11), (17, 5), (10, 3), (3, 1)):
        fine = np.random.default_rng(seed).normal(size=(n, n, n))
        k = np.array([1.0, 2.0, 1.0])

        def conv(a):
            out = a
            for axis in range(3):
                pad = np.pad(out, [(1, 1) if ax == axis else (0, 0) for ax in range(3)])
                sl = lambda s: tuple(slice(s, s + n) if ax == axis else slice(None)
                                     for ax in range(3))
                out = k[0] * pad[sl(0)] + k[1] * pad[sl(1)] + k[2] * pad[sl(2)]
            return out

        expected = (conv(fine) / conv(np.ones_like(fine)))[::2, ::2, ::2]
        assert np.allclose(oracle.restrict(fine), expected, rtol=1e-14, atol=1e-15)


def test_prolongation_known_answers_and_linear_reproduction():
    # reference tests/test_kernels.py:67-72
    coarse = np.arange(125, dtype=np.float64).reshape(5, 5, 5)
    fine = oracle.prolong_add(coarse, np.zeros((9, 9, 9)))
    assert np.array_equal(fine[::2, ::2, ::2], coarse)
    assert fine[1, 1, 1] == pytest.approx(coarse[:2, :2, :2].mean())
    # trilinear interpolation reproduces any trilinear function exactly (dyadic grid)
    c = np.indices((5, 5, 5)).astype(np.float64)
    f = np.indices((9, 9, 9)).astype(np.float64) / 2
    lin = lambda g: 1.0 + 2.0 * g[2] - 4.0 * g[1] + 8.0 * g[0] + 0.5 * g[2] * g[1] * g[0]
    base = np.full((9, 9, 9), 3.0)
    out = oracle.prolong_add(lin(c), base)
    assert np.array_equal(out, base + lin(f))


def test_iso_value_known_answers():
    # reference tests/test_kernels.py:74-88
    coarse = np.arange(125, dtype=np.float64).reshape(5, 5, 5)
    fine = oracle.prolong_add(coarse, np.zeros((9, 9, 9)))
    points = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [1.0, 1.0, 1.0]])
    normals = np.ones_like(points)
    iso = oracle.iso_value(fine, points, normals, False)
    assert iso == pytest.approx(np.mean([fine[0, 0, 0], fine[4, 4, 4], fine[8, 8, 8]]))
    normals[1] *= 2.0
    weighted = oracle.iso_value(fine, points, normals, True)
    expected = np.average([fine[0, 0, 0], fine[4, 4, 4], fine[8, 8, 8]],
                          weights=np.linalg.norm(normals, axis=1))
    assert weighted == pytest.approx(expected)


@pytest.mark.parametrize("n", [9, 65])
def test_residual_closed_form(n):
    # reference tests/test_kernels.py:124-142
    rng = np.random.default_rng(n)
    solution = rng.normal(size=(n, n, n))
    rhs = rng.normal(size=(n, n, n))
    screen = rng.random((n, n, n))
    s, d = stencil_sum_degree(solution)
    h2 = float((n - 1) ** 2)
    expected = rhs - ((d * h2 + 4.0 * screen * h2) * solution - s * h2)
    actual, norm = oracle.residual(solution, rhs, screen, 4.0)
    assert np.allclose(actual, expected, rtol=2e-15, atol=3e-11)
    assert norm == pytest.approx(np.linalg.norm(expected), rel=2e-15)


@pytest.mark.parametrize("n", [2, 9, 65])
def test_red_black_relaxation_closed_form(n):
    # reference tests/test_kernels.py:145-165
    rng = np.random.default_rng(71)
    solution = rng.normal(size=(n, n, n))
    expected = solution.copy()
    rhs = rng.normal(size=(n, n, n))
    screen = rng.random((n, n, n))
    h2 = float((n - 1) ** 2)
    parity = np.indices((n, n, n)).sum(axis=0) & 1
    for color in range(2):
        s, d = stencil_sum_degree(expected)
        diag = d * h2 + 4.0 * screen * h2
        mask = (parity == color) & (diag > 1e-30)
        expected[mask] = (rhs[mask] + s[mask] * h2) / diag[mask]
    actual = oracle.relax(solution, rhs, screen, 4.0, 1)
    assert np.allclose(actual, expected, rtol=0, atol=2e-15)


@pytest.mark.parametrize("n", [9, 33, 65])
@pytest.mark.parametrize("point_weight", [0.0, 4.0])
def test_colour_one_residual_is_rounding_after_relaxation(n, point_weight):
    # Each red-black iteration ends with a colour-1 sweep, which sets every colour-1 node to
    # the value that zeroes its own equation (its diagonal is >= 3 h^-2 > 0 for
    # point_weight >= 0). The colour-1 residual left behind is rounding, a few ulps of the
    # terms it sums, while colour 0's is not: the premise on which the FP32 V-cycle computes
    # and restricts colour 0 only after pre-smoothing.
    rng = np.random.default_rng(n)
    h2 = float((n - 1) ** 2)
    x = rng.normal(size=(n, n, n))
    rhs = rng.normal(size=(n, n, n)) * h2
    screen = rng.random((n, n, n))
    x = oracle.relax(x, rhs, screen, point_weight, 2)
    r, _ = oracle.residual(x, rhs, screen, point_weight)
    s, d = stencil_sum_degree(x)
    terms = np.abs(rhs) + (d * h2 + point_weight * screen * h2) * np.abs(x) + np.abs(s) * h2
    parity = np.indices((n, n, n)).sum(axis=0) & 1
    eps = np.finfo(np.float64).eps
    assert np.all(np.abs(r[parity == 1]) <= 8 * eps * terms[parity == 1])
    assert np.abs(r[parity == 0]).max() > 1e-3 * terms[parity == 0].max()


@pytest.mark.parametrize("n", [14, 65])
def test_dual_vertex_count_closed_form(n):
    # reference tests/test_kernels.py:168-185
    field = np.random.default_rng(n).normal(size=(n, n, n))
    below = field < 0.0
    cnt = sum(below[dz:dz + n - 1, dy:dy + n - 1, dx:dx + n - 1]
              for dz in range(2) for dy in range(2) for dx in range(2))
    assert oracle.dual_vertex_count(field, 0.0) == np.count_nonzero((cnt != 0) & (cnt != 8))


def test_dual_contour_matches_reference_translation_vectors():
    # reference tests/test_kernels.py:188-195
    g = _golden("dual_sphere13.npz")
    v, f, _ = oracle.extract_isosurface(g["field"], float(g["iso"]))
    assert np.allclose(v, g["vertices"], rtol=0, atol=2e-16)
    assert np.array_equal(f, g["faces"])


def test_dual_vertex_capacity_and_cell_map_protocol():
    g = _golden("dual_sphere13.npz")
    field = g["field"]
    total = oracle.dual_vertex_count(field, 0.0)
    cnt, cm, v, d = oracle.dual_vertices(field, field, 0.0, total)
    assert cnt == total and cm.max() == total - 1
    assert np.array_equal(np.sort(cm[cm >= 0]), np.arange(total))  # raster-order indices
    # capacity smaller than the count: counts all, writes the first `cap`, leaves 0 marks
    cap = total // 3
    cnt2, cm2, v2, _ = oracle.dual_vertices(field, field, 0.0, cap)
    assert cnt2 == total
    assert np.array_equal(v2, v[:cap])
    assert np.array_equal(cm2[cm >= cap], np.zeros(np.count_nonzero(cm >= cap), np.int64))
    fcount, _ = oracle.dual_faces(field, 0.0, cm, 0)
    fc2, f2 = oracle.dual_faces(field, 0.0, cm, fcount - 3)  # odd capacity: last quad dropped
    assert fc2 == fcount
    full = oracle.dual_faces(field, 0.0, cm, fcount)[1]
    assert np.array_equal(f2[: fcount - 4], full[: fcount - 4])


def sphere_points(count, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, count)
    theta = rng.uniform(0.0, 2.0 * np.pi, count)
    r = np.sqrt(1.0 - z * z)
    return np.ascontiguousarray(np.column_stack((r * np.cos(theta), r * np.sin(theta), z)))


def test_sphere_reconstruction_invariants():
    # reference tests/test_reconstruction.py:65-83
    points = sphere_points(1200, 2)
    m = oracle.reconstruct(points, points, depth=5, cycles=5)
    f, v = m["faces"], m["vertices"]
    edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    ext = v.max(axis=0) - v.min(axis=0)
    assert ext.max() / ext.min() < 1.04
    vol = abs(np.einsum("ij,ij->i", v[f[:, 0]], np.cross(v[f[:, 1]], v[f[:, 2]])).sum() / 6)
    assert vol == pytest.approx(4 * np.pi / 3, rel=0.04)
    assert m["residual_norm"] / np.linalg.norm(points) < 0.2


def test_vcycle_reduces_residual_monotonically():
    points = sphere_points(800, 9)
    unit, _, _ = oracle.normalize(points, 1.1)
    norms = [oracle.solve(unit, points, 5, cycles=c)["residual_norm"] for c in (1, 2, 4, 8)]
    assert all(a > b for a, b in zip(norms, norms[1:]))


def test_empty_solve_is_zero():
    sol = oracle.solve(np.empty((0, 3)), np.empty((0, 3)), 3)
    assert sol["residual_norm"] == 0.0 and sol["iso_value"] == 0.0
    assert not sol["field"].any()


@pytest.mark.parametrize("n", [2, 3, 5, 9, 10])
def test_numpy_smoother_restatement_equals_c_oracle_in_float64(n):
    # pins relax_np / residual_np, whose float32 instantiation checks the FP32 kernels
    rng = np.random.default_rng(50 + n)
    x, rhs = rng.normal(size=(n, n, n)), rng.normal(size=(n, n, n))
    screen = np.where(rng.random((n, n, n)) < 0.3, rng.random((n, n, n)), 0.0)
    assert np.array_equal(oracle.relax_np(x, rhs, screen, 4.0, 3),
                          oracle.relax(x, rhs, screen, 4.0, 3))
    assert np.array_equal(oracle.residual_np(x, rhs, screen, 4.0),
                          oracle.residual(x, rhs, screen, 4.0)[0])
    x32, r32, s32 = x.astype(np.float32), rhs.astype(np.float32), screen.astype(np.float32)
    got = oracle.relax_np(x32, r32, s32, 4.0, 3)
    assert got.dtype == np.float32
    want = oracle.relax(x32.astype(np.float64), r32.astype(np.float64), s32.astype(np.float64),
                        4.0, 3)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-5 * np.abs(want).max())


@pytest.mark.parametrize("n", [2, 3, 5, 9, 10, 17])
def test_numpy_transfer_restatements_equal_c_oracle_in_float64(n):
    # pins restrict_np / prolong_np, whose float32 instantiation checks the FP32 kernels
    rng = np.random.default_rng(90 + n)
    fine = rng.normal(size=(n, n, n))
    cn = (n + 1) // 2
    coarse = rng.normal(size=(cn, cn, cn))
    assert np.array_equal(oracle.restrict_np(fine), oracle.restrict(fine))
    assert np.array_equal(oracle.restrict_np(fine, 4.0), 4.0 * oracle.restrict(fine))
    assert np.array_equal(oracle.prolong_np(coarse, fine), oracle.prolong_add(coarse, fine))
