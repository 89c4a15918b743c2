"""PLY output of reconstructed meshes (the format upstream PoissonRecon writes; the
reference package returns arrays only, README.md:31-35).

``write_ply`` writes a Mesh as binary little-endian (default) or ASCII PLY: per vertex
``double x, y, z`` and, optionally, ``double density`` (the mesh's vertex density, as
upstream's --density output), and per face a ``uchar`` count with ``int`` vertex indices.
``read_ply`` reads those files back (these two layouts only).
"""

from __future__ import annotations

import os

import numpy as np

_VERTEX = [("x", "<f8"), ("y", "<f8"), ("z", "<f8")]
_FACE = np.dtype([("n", "u1"), ("i", "<i4", (3,))])


def write_ply(path: str | os.PathLike, mesh, *, binary: bool = True, density: bool = True) -> None:
    """Write ``mesh`` (a Mesh, or anything with vertices, faces and density) to ``path``."""
    vertices = np.ascontiguousarray(mesh.vertices, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(mesh.faces, dtype=np.int64).reshape(-1, 3)
    if len(faces) and (faces.min() < 0 or faces.max() >= len(vertices)):
        raise ValueError("face index out of range")
    if len(vertices) >= 2 ** 31:
        raise ValueError("PLY int vertex indices hold at most 2^31 - 1 vertices")
    fields = list(_VERTEX) + ([("density", "<f8")] if density else [])
    vrec = np.empty(len(vertices), dtype=np.dtype(fields))
    vrec["x"], vrec["y"], vrec["z"] = vertices[:, 0], vertices[:, 1], vertices[:, 2]
    if density:
        vrec["density"] = np.asarray(mesh.density, dtype=np.float64).reshape(-1)
    header = ["ply", f"format {'binary_little_endian' if binary else 'ascii'} 1.0",
              "comment screened Poisson reconstruction (mojo_poissonrecon)",
              f"element vertex {len(vertices)}"]
    header += [f"property double {name}" for name, _ in fields]
    header += [f"element face {len(faces)}", "property list uchar int vertex_indices",
               "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(header) + "\n").encode("ascii"))
        if binary:
            fh.write(vrec.tobytes())
            frec = np.empty(len(faces), dtype=_FACE)
            frec["n"] = 3
            frec["i"] = faces
            fh.write(frec.tobytes())
        else:
            cols = np.column_stack([vrec[name] for name, _ in fields]) if len(vrec) else \
                np.empty((0, len(fields)))
            np.savetxt(fh, cols, fmt="%.17g")
            np.savetxt(fh, np.column_stack([np.full(len(faces), 3), faces]) if len(faces)
                       else np.empty((0, 4)), fmt="%d")


def read_ply(path: str | os.PathLike):
    """Read a file written by write_ply: returns (vertices, faces, density or None)."""
    with open(path, "rb") as fh:
        lines = []
        while True:
            line = fh.readline()
            if not line:
                raise ValueError("PLY header has no end_header")
            lines.append(line.decode("ascii").strip())
            if lines[-1] == "end_header":
                break
        if lines[0] != "ply":
            raise ValueError("not a PLY file")
        fmt = lines[1].split()[1]
        nv = nf = 0
        names, element = [], None
        for line in lines[2:]:
            parts = line.split()
            if parts[0] == "element":
                element = parts[1]
                if element == "vertex":
                    nv = int(parts[2])
                elif element == "face":
                    nf = int(parts[2])
            elif parts[0] == "property" and element == "vertex":
                names.append(parts[-1])
        vdtype = np.dtype([(name, "<f8") for name in names])
        if fmt == "binary_little_endian":
            vrec = np.frombuffer(fh.read(vdtype.itemsize * nv), dtype=vdtype, count=nv)
            frec = np.frombuffer(fh.read(_FACE.itemsize * nf), dtype=_FACE, count=nf)
            if nf and not np.all(frec["n"] == 3):
                raise ValueError("only triangle faces are supported")
            faces = frec["i"].astype(np.int64)
            cols = np.column_stack([vrec[name] for name in names]) if nv else np.empty((0, len(names)))
        elif fmt == "ascii":
            rows = fh.read().decode("ascii").split("\n")
            cols = np.array([[float(x) for x in r.split()] for r in rows[:nv]]).reshape(nv, len(names))
            frows = np.array([[int(x) for x in r.split()] for r in rows[nv:nv + nf]],
                             dtype=np.int64).reshape(nf, 4)
            faces = frows[:, 1:]
        else:
            raise ValueError(f"unsupported PLY format {fmt}")
    vertices = np.ascontiguousarray(cols[:, :3])
    density = np.ascontiguousarray(cols[:, names.index("density")]) if "density" in names else None
    return vertices, np.ascontiguousarray(faces).reshape(-1, 3), density
