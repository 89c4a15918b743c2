"""Mesh accuracy metrics, computed on the GPU.

The reference judges a reconstruction by a symmetric distance between meshes divided by
the bounding-box diagonal (tests/test_reconstruction.py:50-62). ``mesh_distance`` computes,
for every vertex of each mesh, the exact distance to the nearest point of the other mesh's
triangles (``mpr_mesh_distance`` in include/mpr_hip.h); the Hausdorff distance is the larger
of the two directions' maxima, the vertex chamfer the mean of the two directions' means.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, f64, i64, lib, ptr


@dataclass(frozen=True)
class MeshDistance:
    a_to_b_max: float
    b_to_a_max: float
    a_to_b_mean: float
    b_to_a_mean: float

    @property
    def hausdorff(self) -> float:
        return max(self.a_to_b_max, self.b_to_a_max)

    @property
    def chamfer(self) -> float:
        return 0.5 * (self.a_to_b_mean + self.b_to_a_mean)


def _mesh(m):
    vertices, faces = (m.vertices, m.faces) if hasattr(m, "vertices") else m
    vertices = f64(vertices).reshape(-1, 3)
    faces = i64(faces).reshape(-1, 3)
    return vertices, faces


def mesh_distance(a, b) -> MeshDistance:
    """Distances between two triangle meshes (``Mesh`` objects or ``(vertices, faces)``)."""
    va, fa = _mesh(a)
    vb, fb = _mesh(b)
    out = np.zeros(4)
    check(lib().mpr_mesh_distance(ptr(va), len(va), ptr(fa), len(fa), ptr(vb), len(vb),
                                  ptr(fb), len(fb), out.ctypes.data_as(ctypes.c_void_p)),
          "mpr_mesh_distance")
    return MeshDistance(*(float(x) for x in out))


def hausdorff(a, b) -> float:
    """Symmetric vertex-to-surface Hausdorff distance between two meshes."""
    return mesh_distance(a, b).hausdorff
