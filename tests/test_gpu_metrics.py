"""GPU test of the mesh distance evaluator (mpr_mesh_distance, mojo_poissonrecon.metrics).

Checked against an independent CPU computation of the point-to-triangle distance: the
unconstrained nearest point of the triangle's plane (normal equations of the barycentric
least-squares problem) when it lies inside the triangle, else the nearest of the three
edges (the constrained minimum of a convex quadratic lies on the boundary) - a different
method from the GPU's Voronoi-region test. Also: a mesh is at distance 0 from itself, the
distance is bounded by the nearest-vertex distance, and a rigid shift moves a closed mesh by
at most the shift.
"""

from __future__ import annotations

import numpy as np
import pytest

import mojo_poissonrecon as mpr
from mojo_poissonrecon import shapes
from mojo_poissonrecon._lib import DeviceError
from mojo_poissonrecon.metrics import mesh_distance

pytestmark = pytest.mark.gpu


def _seg_dist(p, a, b):
    ab = b - a
    len2 = np.einsum("...i,...i->...", ab, ab)
    t = np.where(len2 > 0, np.einsum("...i,...i->...", p - a, ab) / np.where(len2 > 0, len2, 1), 0)
    t = np.clip(t, 0.0, 1.0)
    return np.linalg.norm(p - (a + t[..., None] * ab), axis=-1)


def brute_distance(points, vertices, faces):
    """distance from each point to the nearest triangle, over every (point, triangle) pair"""
    a, b, c = (vertices[faces[:, k]][None, :, :] for k in range(3))
    p = points[:, None, :]
    e1, e2, w = b - a, c - a, p - a
    m11 = np.einsum("...i,...i->...", e1, e1)
    m12 = np.einsum("...i,...i->...", e1, e2)
    m22 = np.einsum("...i,...i->...", e2, e2)
    r1 = np.einsum("...i,...i->...", e1, w)
    r2 = np.einsum("...i,...i->...", e2, w)
    det = m11 * m22 - m12 * m12
    ok = det > 1e-300
    sdet = np.where(ok, det, 1.0)
    s = (m22 * r1 - m12 * r2) / sdet
    t = (m11 * r2 - m12 * r1) / sdet
    inside = ok & (s >= 0) & (t >= 0) & (s + t <= 1)
    plane = np.linalg.norm(w - s[..., None] * e1 - t[..., None] * e2, axis=-1)
    edges = np.minimum(np.minimum(_seg_dist(p, a, b), _seg_dist(p, b, c)), _seg_dist(p, c, a))
    return np.where(inside, np.minimum(plane, edges), edges).min(axis=1)


def random_mesh(rng, nv, nf, scale=1.0):
    v = rng.normal(size=(nv, 3)) * scale
    f = np.stack([rng.choice(nv, 3, replace=False) for _ in range(nf)])
    return v, f


def test_distance_equals_brute_force_on_random_meshes():
    rng = np.random.default_rng(5)
    for nv_a, nf_a, nv_b, nf_b, shift in ((40, 60, 50, 90, 0.0), (300, 500, 200, 400, 0.3),
                                          (7, 1, 500, 900, 5.0)):
        va, fa = random_mesh(rng, nv_a, nf_a)
        vb, fb = random_mesh(rng, nv_b, nf_b)
        vb = vb + shift
        d = mesh_distance((va, fa), (vb, fb))
        ab = brute_distance(va, vb, fb)
        ba = brute_distance(vb, va, fa)
        assert d.a_to_b_max == pytest.approx(ab.max(), rel=1e-12, abs=1e-15)
        assert d.b_to_a_max == pytest.approx(ba.max(), rel=1e-12, abs=1e-15)
        assert d.a_to_b_mean == pytest.approx(ab.mean(), rel=1e-12)
        assert d.b_to_a_mean == pytest.approx(ba.mean(), rel=1e-12)


def test_distance_on_reconstructed_meshes():
    from scipy.spatial import cKDTree

    p, n = shapes.sphere(4000, seed=3)
    m = mpr.reconstruct(p, n, depth=6)
    d0 = mesh_distance(m, m)
    assert d0.hausdorff == 0.0 and d0.chamfer == 0.0
    shift = np.array([1e-3, -2e-3, 0.5e-3])
    moved = (m.vertices + shift, m.faces)
    d = mesh_distance(m, moved)
    assert 0.0 < d.hausdorff <= np.linalg.norm(shift) * (1 + 1e-12)
    vertex_bound = max(cKDTree(moved[0]).query(m.vertices)[0].max(),
                       cKDTree(m.vertices).query(moved[0])[0].max())
    assert d.hausdorff <= vertex_bound * (1 + 1e-12)
    sample = np.random.default_rng(0).choice(len(m.vertices), 100, replace=False)
    want = brute_distance(m.vertices[sample], moved[0], m.faces).max()
    assert d.a_to_b_max >= want * (1 - 1e-12)
    # a different depth of the same cloud: the GPU value is the brute force's on a sample
    m5 = mpr.reconstruct(p, n, depth=5)
    d5 = mesh_distance(m, m5)
    got = brute_distance(m.vertices[sample], m5.vertices, m5.faces)
    assert got.max() <= d5.a_to_b_max * (1 + 1e-12)


def test_distance_edge_cases():
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    f = np.array([[0, 1, 2]])
    q = (np.array([[0.25, 0.25, 2.0]]), np.empty((0, 3), np.int64))
    d = mesh_distance(q, (v, f))
    assert d.a_to_b_max == pytest.approx(2.0, rel=1e-15)
    assert d.b_to_a_max == np.inf  # vertices but no triangle to measure against
    empty = (np.empty((0, 3)), np.empty((0, 3), np.int64))
    d = mesh_distance(empty, empty)
    assert d.hausdorff == 0.0
    # a degenerate (zero-area) triangle measures as its edges
    line = (np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]]), np.array([[0, 1, 2]]))
    d = mesh_distance((np.array([[1.0, 3.0, 4.0]]), np.empty((0, 3), np.int64)), line)
    assert d.a_to_b_max == pytest.approx(5.0, rel=1e-15)
    with pytest.raises(DeviceError, match="out of range"):
        mesh_distance((v, np.array([[0, 1, 3]])), (v, f))
