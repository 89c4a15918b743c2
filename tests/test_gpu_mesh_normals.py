"""GPU parity of oriented_points_from_mesh (core.py:429-450) against the oracle's per-face
loop: bit for bit on the reference's degenerate-mesh fixtures
(tests/test_reconstruction.py:86-179 of the reference), on random meshes with degenerate,
huge and tiny faces, and on a reconstructed mesh of 1M+ faces; and reconstruct_mesh on top of
it equals reconstruct() of the oracle's points.
"""

from __future__ import annotations

import numpy as np
import pytest

import mojo_poissonrecon as mpr
import oracle
from mojo_poissonrecon import shapes

from test_boundary import (DUPLICATE_VERTICES, EMPTY, NON_MANIFOLD, SINGLE_TRIANGLE,
                           UNREFERENCED, ZERO_AREA)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("vertices,faces,expected",
                         [EMPTY, SINGLE_TRIANGLE, DUPLICATE_VERTICES, ZERO_AREA, NON_MANIFOLD,
                          UNREFERENCED])
def test_reference_mesh_fixtures_bit_exact(vertices, faces, expected):
    points, normals = mpr.oriented_points_from_mesh(vertices, faces)
    want_p, want_n = oracle.oriented_points_from_mesh(vertices, faces)
    assert len(points) == expected
    assert np.array_equal(points, want_p) and np.array_equal(normals, want_n)


def test_area_weighted_known_answer():
    vertices = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [0, 0, 1]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [0, 3, 1]], dtype=np.int64)
    points, normals = mpr.oriented_points_from_mesh(vertices, faces)
    origin = np.flatnonzero(np.all(points == vertices[0], axis=1))[0]
    assert np.allclose(normals[origin], [0.0, 1 / np.sqrt(5), 2 / np.sqrt(5)])


def test_random_meshes_bit_exact():
    rng = np.random.default_rng(5)
    for nv, nf, mag in ((300, 1000, 0), (20_000, 80_000, 150)):
        vertices = rng.normal(size=(nv, 3)) * 10.0 ** rng.integers(-mag, mag + 1, size=(nv, 1))
        faces = rng.integers(0, nv, size=(nf, 3))
        faces[:10, 1] = faces[:10, 0]  # degenerate faces
        vertices[5] = np.nan  # faces through it are skipped
        with np.errstate(all="ignore"):
            want_p, want_n = oracle.oriented_points_from_mesh_c(vertices, faces)
        got_p, got_n = mpr.oriented_points_from_mesh(vertices, faces)
        assert np.array_equal(got_p, want_p) and np.array_equal(got_n, want_n)


def test_million_face_mesh_and_reconstruct_mesh():
    p, n = shapes.torus_workload(400_000, seed=2)
    mesh = mpr.reconstruct(p, n, depth=9, confidence=True, precision="mixed")
    assert len(mesh.faces) > 1_000_000
    print(f"{len(mesh.vertices)} vertices, {len(mesh.faces)} faces", flush=True)
    got_p, got_n = mpr.oriented_points_from_mesh(mesh.vertices, mesh.faces)
    want_p, want_n = oracle.oriented_points_from_mesh_c(mesh.vertices, mesh.faces)
    assert np.array_equal(got_p, want_p) and np.array_equal(got_n, want_n)
    # reconstruct_mesh = reconstruct of those oriented points
    small = mpr.reconstruct(p[:20_000], n[:20_000], depth=6)
    again = mpr.reconstruct_mesh(small.vertices, small.faces, depth=6)
    sp, sn = oracle.oriented_points_from_mesh_c(small.vertices, small.faces)
    want = mpr.reconstruct(sp, sn, depth=6)
    assert np.array_equal(again.faces, want.faces) and np.array_equal(again.vertices, want.vertices)
