"""Write digests of the oracle's reconstruction of the 2M-point depth-9 superquadric.

The oracle (the serial C restatement of the reference, oracle/) takes about a minute for
this workload, too long to run beside every GPU test, so its result is recorded once here
and the GPU test compares against the record. The input is regenerated from its seed by
mojo_poissonrecon.shapes (the same NumPy on both sides). Recorded: the SHA-256 of the face
array (int64, C order), the vertex and face counts, the iso value and residual norm, every
1000th vertex and vertex density, and the boundary edges of the mesh (edges used by one
face), which this input has where the surface reaches the grid's boundary planes.

Run from the repository root: python tests/golden/make_oracle_digests.py
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mojo-poissonrecon_amd"))

import oracle  # noqa: E402
from mojo_poissonrecon import shapes  # noqa: E402

OUT = os.path.join(HERE, "superquadric2m_depth9_fp64.json")


def face_digest(faces: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(faces, dtype=np.int64).tobytes()).hexdigest()


def edge_counts(faces: np.ndarray):
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * (int(e.max()) + 1) + e[:, 1]
    return np.unique(key, return_counts=True)


def main() -> None:
    p, n = shapes.superquadric_workload(2_000_000, seed=0)
    want = oracle.reconstruct(p, n, depth=9)
    v, f, d = want["vertices"], want["faces"], want["density"]
    _, counts = edge_counts(f)
    record = {
        "workload": "shapes.superquadric_workload(2_000_000, seed=0), depth 9, FP64, defaults",
        "vertices": int(len(v)),
        "faces": int(len(f)),
        "faces_sha256": face_digest(f),
        "iso_value": float(want["iso_value"]),
        "residual_norm": float(want["residual_norm"]),
        "edges_by_use": {str(int(k)): int(c) for k, c in zip(*np.unique(counts, return_counts=True))},
        "vertex_sample_stride": 1000,
        "vertex_sample": v[::1000].tolist(),
        "density_sample": d[::1000].tolist(),
    }
    with open(OUT, "w") as fh:
        json.dump(record, fh)
    print(OUT, {k: record[k] for k in ("vertices", "faces", "iso_value", "edges_by_use")})


if __name__ == "__main__":
    main()
