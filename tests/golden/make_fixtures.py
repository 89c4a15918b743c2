"""Write the golden vectors under tests/golden/ from the reference's NumPy translations.

Run once in a checkout that has the reference at /root/reference (it never runs in the
test suite and the reference does not travel). It loads
``python/mojo_poissonrecon/_reference.py`` by file path (that module imports only
NumPy) and records inputs and outputs for the two functions whose details no closed
form in the reference tests pins completely:

* ``dual_isosurface`` on the 13^3 sphere distance field of
  tests/test_kernels.py:188-195 (vertex positions and the exact face order);
* ``splat`` + ``assemble`` on the three hand-placed points of
  tests/test_kernels.py:24-43 (the density-normalised vector field and the RHS).

The fixtures are data only: arrays in .npz files, loaded with allow_pickle=False.
"""

from __future__ import annotations

import importlib.util
import os

import numpy as np

REF = "/root/reference/python/mojo_poissonrecon/_reference.py"
OUT = os.path.dirname(os.path.abspath(__file__))


def load_reference():
    spec = importlib.util.spec_from_file_location("_mpr_reference_translation", REF)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def main() -> None:
    ref = load_reference()

    # tests/test_kernels.py:188-195
    axis = np.linspace(0.0, 1.0, 13)
    z, y, x = np.meshgrid(axis, axis, axis, indexing="ij")
    field = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 - 0.3**2
    vertices, faces = ref.dual_isosurface(field, 0.0)
    np.savez(os.path.join(OUT, "dual_sphere13.npz"), field=field, iso=np.float64(0.0),
             vertices=np.asarray(vertices, np.float64), faces=np.asarray(faces, np.int64))

    # tests/test_kernels.py:24-43
    points = np.array([[0.15, 0.25, 0.35], [0.8, 0.7, 0.6], [1.0, 0.5, 0.0]], dtype=np.float64)
    normals = np.array([[2.0, 0.0, 0.0], [0.0, -3.0, 0.0], [0.0, 0.0, 4.0]])
    density, vx, vy, vz, screen = ref.splat(points, normals, 5, True, samples_per_node=1.0)
    rhs = ref.assemble(vx, vy, vz, screen, 4.0)
    np.savez(os.path.join(OUT, "splat3_depth2.npz"), points=points, normals=normals,
             n=np.int64(5), confidence=np.int64(1), samples_per_node=np.float64(1.0),
             point_weight=np.float64(4.0), density=density, vx=vx, vy=vy, vz=vz,
             screen=screen, rhs=rhs)
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
