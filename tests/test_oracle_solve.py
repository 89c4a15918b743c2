"""CPU checks that pin the oracle's solver orchestration (oracle_solve, restating
core.py:196-306 of the reference) without retyping it.

Three independent anchors:

1. Known answer of the density normalisation: the trilinear weights of one sample sum to
   one, so the splat adds 2 w per sample (capi.mojo:127-173) and core.py:255 divides by
   samples_per_node: sum(density) = 2 sum(w) / spn, sum(screen) = sum(w), for spn != 1.
2. The fixed point: with enough V-cycles the solution is the solution of the discrete
   screened Poisson system, built here from the reference's closed-form stencil
   (tests/test_kernels.py:106-121 of the reference: Neumann 7-point neighbour sum and
   degree) with diagonal deg h^-2 + alpha s h^-2 and the assembled right-hand side, and
   solved directly (scipy sparse LU). This pins the fine operator, the rhs target 0.5 and
   convergence, whatever the coarse levels do.
3. One cycle, bit for bit: oracle_solve(cycles=1) equals the composition, in core.py's
   order (screen pyramid x4, initial guess 0.5, coarse x = 0 before each descent, pre 3 /
   post 3, 40 iterations on the 3^3 grid, restriction of the residual, prolongation before
   post-smoothing), of the NumPy kernel restatements relax_np, residual_np, restrict_np and
   prolong_np, which are pinned on their own (bit-equal to the C kernels, which match the
   reference's closed forms).

The last test builds the oracle with each of nine deliberate slips in the orchestration
(oracle/mpr_oracle.c, ORACLE_MUTATION) and shows that at least one anchor rejects every one.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle


def _cloud(count: int, seed: int):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    points = 0.5 + 0.3 * d + rng.normal(0, 0.01, d.shape)
    normals = d * rng.uniform(0.5, 1.5, (count, 1))
    return np.ascontiguousarray(points), np.ascontiguousarray(normals)


# ---- anchor 1: density known answer ----------------------------------------------------

def check_density_known_answer(solve) -> None:
    points, normals = _cloud(40, 5)
    for spn, confidence in ((2.5, False), (0.75, True)):
        got = solve(points, normals, 3, confidence=confidence, cycles=1, samples_per_node=spn)
        w = np.linalg.norm(normals, axis=1) if confidence else np.ones(len(points))
        assert got["density"].sum() == pytest.approx(2.0 * w.sum() / spn, rel=1e-13)
        assert got["screen"].sum() == pytest.approx(w.sum(), rel=1e-13)


# ---- anchor 2: the fixed point is the discrete system's solution -----------------------

def _stencil_sum_degree(values):
    """The reference's closed form (tests/test_kernels.py:106-121)."""
    neighbor_sum = np.zeros_like(values)
    degree = np.zeros(values.shape, dtype=np.int64)
    neighbor_sum[:, :, 1:] += values[:, :, :-1]
    degree[:, :, 1:] += 1
    neighbor_sum[:, :, :-1] += values[:, :, 1:]
    degree[:, :, :-1] += 1
    neighbor_sum[:, 1:, :] += values[:, :-1, :]
    degree[:, 1:, :] += 1
    neighbor_sum[:, :-1, :] += values[:, 1:, :]
    degree[:, :-1, :] += 1
    neighbor_sum[1:, :, :] += values[:-1, :, :]
    degree[1:, :, :] += 1
    neighbor_sum[:-1, :, :] += values[1:, :, :]
    degree[:-1, :, :] += 1
    return neighbor_sum, degree


def discrete_solution(points, normals, depth, point_weight, confidence, spn):
    """Solve (deg h^-2 + alpha s h^-2) x - h^-2 sum_nbr x = rhs directly."""
    n = (1 << depth) + 1
    _, _, vx, vy, vz, screen = oracle.splat(points, normals, n, confidence, spn)
    rhs = oracle.assemble(vx, vy, vz, screen, point_weight, 0.5)
    h2 = float((n - 1) ** 2)
    m = n ** 3
    _, degree = _stencil_sum_degree(np.zeros((n, n, n)))
    idx = np.arange(m).reshape(n, n, n)
    rows, cols = [], []
    for a, b in ((idx[:, :, 1:], idx[:, :, :-1]), (idx[:, 1:, :], idx[:, :-1, :]),
                 (idx[1:, :, :], idx[:-1, :, :])):
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
    adjacency = sp.csr_matrix((np.ones(sum(len(r) for r in rows)),
                               (np.concatenate(rows), np.concatenate(cols))), shape=(m, m))
    diag = degree.ravel() * h2 + point_weight * screen.ravel() * h2
    A = sp.diags(diag) - h2 * adjacency
    # the discrete operator of the closed form applied to a random vector, as a check of A
    probe = np.random.default_rng(0).normal(size=(n, n, n))
    s, deg = _stencil_sum_degree(probe)
    applied = (deg * h2 + point_weight * screen * h2) * probe - s * h2
    assert np.allclose(A @ probe.ravel(), applied.ravel(), rtol=1e-13, atol=1e-9)
    return spla.spsolve(A.tocsc(), rhs.ravel()).reshape(n, n, n)


def check_fixed_point(solve) -> None:
    points, normals = _cloud(60, 7)
    for depth, pw, confidence in ((2, 4.0, True), (3, 4.0, False), (3, 25.0, True)):
        want = discrete_solution(points, normals, depth, pw, confidence, 1.5)
        got = solve(points, normals, depth, point_weight=pw, confidence=confidence, cycles=60)
        scale = np.abs(want).max()
        assert np.abs(got["field"] - want).max() <= 1e-10 * scale, (depth, pw)


# ---- anchor 3: one cycle is the composition of the pinned kernels -----------------------

def composed_solve(points, normals, depth, point_weight=4.0, confidence=False, cycles=1,
                   samples_per_node=1.5):
    """core.py:229-306 written with the NumPy kernel restatements."""
    n = (1 << depth) + 1
    _, density, vx, vy, vz, screen = oracle.splat(points, normals, n, confidence,
                                                  samples_per_node)
    density = density / samples_per_node
    rhs = oracle.assemble(vx, vy, vz, screen, point_weight, 0.5)
    screens = [screen]
    while screens[-1].shape[0] > 3:
        screens.append(oracle.restrict_np(screens[-1]) * 4.0)
    solutions = [np.full_like(s, 0.5) for s in screens]
    right = [np.zeros_like(s) for s in screens]
    right[0][...] = rhs

    def v_cycle(level):
        x, b, s = solutions[level], right[level], screens[level]
        if x.shape[0] <= 3:
            solutions[level] = oracle.relax_np(x, b, s, point_weight, 40)
            return
        x = oracle.relax_np(x, b, s, point_weight, 3)
        r = oracle.residual_np(x, b, s, point_weight)
        right[level + 1] = oracle.restrict_np(r)
        solutions[level + 1] = np.zeros_like(solutions[level + 1])
        v_cycle(level + 1)
        x = oracle.prolong_np(solutions[level + 1], x)
        solutions[level] = oracle.relax_np(x, b, s, point_weight, 3)

    for _ in range(cycles):
        v_cycle(0)
    return dict(field=solutions[0], density=density, screen=screen, rhs=rhs)


def check_one_cycle_composition(solve) -> None:
    points, normals = _cloud(50, 11)
    for depth, cycles, spn in ((2, 1, 1.5), (3, 1, 2.0), (3, 2, 1.5)):
        got = solve(points, normals, depth, confidence=True, cycles=cycles, samples_per_node=spn)
        want = composed_solve(points, normals, depth, confidence=True, cycles=cycles,
                              samples_per_node=spn)
        for name in ("density", "screen", "rhs", "field"):
            assert np.array_equal(got[name], want[name]), (name, depth, cycles)


ANCHORS = (check_density_known_answer, check_fixed_point, check_one_cycle_composition)


@pytest.mark.parametrize("anchor", ANCHORS, ids=lambda f: f.__name__)
def test_oracle_solve_is_pinned(anchor):
    anchor(oracle.solve)


MUTATIONS = {
    1: "screen pyramid x2 instead of x4",
    2: "initial guess 0 instead of 0.5",
    3: "coarse solution not zeroed before the descent",
    4: "post-smoothing before the prolongation",
    5: "density not divided by samples_per_node",
    6: "2 pre-smoothing iterations instead of 3",
    7: "20 coarsest iterations instead of 40",
    8: "rhs target 1.0 instead of 0.5",
    9: "restriction of the solution instead of the residual",
}


@pytest.mark.parametrize("mutation", sorted(MUTATIONS), ids=lambda k: MUTATIONS[k])
def test_every_orchestration_slip_fails_an_anchor(mutation, tmp_path):
    handle = oracle.build_mutation(mutation, str(tmp_path))

    def mutated(*args, **kwargs):
        return oracle.solve(*args, handle=handle, **kwargs)

    failed = []
    for anchor in ANCHORS:
        try:
            anchor(mutated)
        except AssertionError:
            failed.append(anchor.__name__)
    assert failed, f"no anchor rejects: {MUTATIONS[mutation]}"


def test_openmp_oracle_equals_the_serial_oracle():
    # the all-cores build (timed as the CPU baseline) computes the same grids bit for bit
    points, normals = _cloud(3000, 13)
    omp = oracle.openmp()
    assert omp.oracle_threads() >= 1
    for depth in (4, 6):
        a = oracle.solve(points, normals, depth, confidence=True)
        b = oracle.solve(points, normals, depth, confidence=True, handle=omp)
        for name in ("density", "screen", "rhs", "field"):
            assert np.array_equal(a[name], b[name]), (name, depth)
        assert a["residual_norm"] == b["residual_norm"] and a["iso_value"] == b["iso_value"]
