"""CPU oracle for the screened Poisson hot path. TEST INFRASTRUCTURE ONLY.

This package is the checker: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import it; the product package never does. It is a
serial FP64 restatement of the reference:

* ``mpr_oracle.c`` restates the kernels of ``src/capi.mojo`` op for op and the solver
  orchestration of ``python/mojo_poissonrecon/core.py:196-306``;
* this module restates the NumPy-side logic of ``core.py``: ``build_octree``
  (core.py:86-127), ``_valid_oriented_rows`` (core.py:73-83), ``_normalize``
  (core.py:377-385), ``extract_isosurface``'s count/fill protocol (core.py:309-374),
  ``reconstruct`` (core.py:388-426) and ``oriented_points_from_mesh`` (core.py:429-450).

Validation error messages are not restated here; the product package owns them.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mpr_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_I = ctypes.c_int64
_F = ctypes.c_double

_SIGNATURES = {
    "oracle_norm3": ([_F, _F, _F], _F),
    "oracle_threads": ([], ctypes.c_int),
    "oracle_octree_codes": ([_P, _I, _I, _P], None),
    "oracle_splat": ([_P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _F], _I),
    "oracle_assemble": ([_P, _P, _P, _P, _P, _I, _F, _F], None),
    "oracle_relax": ([_P, _P, _P, _I, _F, _I], None),
    "oracle_residual": ([_P, _P, _P, _P, _I, _F], _F),
    "oracle_restrict": ([_P, _P, _I], None),
    "oracle_prolong_add": ([_P, _P, _I], None),
    "oracle_iso_value": ([_P, _P, _P, _I, _I, _I], _F),
    "oracle_dual_vertices": ([_P, _P, _I, _F, _P, _P, _P, _I], _I),
    "oracle_dual_faces": ([_P, _I, _F, _P, _P, _I], _I),
    "oracle_solve": ([_P, _P, _I, _I, _F, _I, _I, _F, _P, _P, _P, _P, _P, _P], _I),
    "oracle_oriented_points": ([_P, _I, _P, _I, _P, _P], _I),
}


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction, IEEE double arithmetic)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= max(
        os.path.getmtime(SRC), os.path.getmtime(os.path.join(HERE, "mpr_oracle.h"))
    ):
        return LIB
    tmp = LIB + ".tmp"
    cmd = [
        "gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
        "-fno-fast-math", SRC, "-o", tmp, "-lm",
    ]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def _load(path: str) -> ctypes.CDLL:
    handle = ctypes.CDLL(path)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = res
    return handle


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = _load(build())
    return _lib


OMP_LIB = os.path.join(HERE, "liboracle_omp.so")
_omp = None


def openmp() -> ctypes.CDLL:
    """The oracle built with -fopenmp (its plane loops on all host threads, results equal
    to the serial build's), for timing the reference algorithm on every core."""
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_LIB) or os.path.getmtime(OMP_LIB) < os.path.getmtime(SRC):
            tmp = OMP_LIB + ".tmp"
            cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                   "-fno-fast-math", "-fopenmp", SRC, "-o", tmp, "-lm"]
            subprocess.run(cmd, check=True, capture_output=True, text=True)
            os.replace(tmp, OMP_LIB)
        _omp = _load(OMP_LIB)
    return _omp


def build_mutation(mutation: int, out_dir: str) -> ctypes.CDLL:
    """A copy of the oracle built with -DORACLE_MUTATION=`mutation` (a deliberate slip in
    the solver orchestration, mpr_oracle.c) into out_dir, for tests that show their checks
    reject it. Never used as the checker."""
    path = os.path.join(out_dir, f"liboracle_mutation{int(mutation)}.so")
    cmd = [
        "gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
        f"-DORACLE_MUTATION={int(mutation)}", SRC, "-o", path, "-lm",
    ]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return _load(path)


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.int64), a.dtype
    return a.ctypes.data if a.size else 0


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- kernel-level restatements (capi.mojo) -------------------------------------------

def norm3(x: float, y: float, z: float) -> float:
    return lib().oracle_norm3(x, y, z)


def octree_codes(points, depth: int) -> np.ndarray:
    points = _f(points)
    codes = np.empty((len(points), depth + 1), dtype=np.int64)
    lib().oracle_octree_codes(_p(points), len(points), depth, _p(codes))
    return codes


def splat(points, normals, n: int, confidence: bool, samples_per_node: float = 1.5):
    """Returns (accepted, density, vx, vy, vz, screen); density is the raw 2w sum."""
    points, normals = _f(points), _f(normals)
    grids = [np.zeros((n, n, n)) for _ in range(5)]
    density, vx, vy, vz, screen = grids
    acc = lib().oracle_splat(_p(points), _p(normals), len(points), n, int(confidence),
                             _p(density), _p(vx), _p(vy), _p(vz), _p(screen),
                             float(samples_per_node))
    return acc, density, vx, vy, vz, screen


def assemble(vx, vy, vz, screen, point_weight: float, target: float = 0.5) -> np.ndarray:
    vx, vy, vz, screen = map(_f, (vx, vy, vz, screen))
    rhs = np.empty_like(vx)
    lib().oracle_assemble(_p(vx), _p(vy), _p(vz), _p(screen), _p(rhs), vx.shape[0],
                          float(point_weight), float(target))
    return rhs


def relax(solution, rhs, screen, point_weight: float, iterations: int) -> np.ndarray:
    out = np.array(solution, dtype=np.float64, copy=True, order="C")
    rhs, screen = _f(rhs), _f(screen)
    lib().oracle_relax(_p(out), _p(rhs), _p(screen), out.shape[0], float(point_weight),
                       int(iterations))
    return out


def residual(solution, rhs, screen, point_weight: float):
    solution, rhs, screen = map(_f, (solution, rhs, screen))
    out = np.empty_like(solution)
    norm = lib().oracle_residual(_p(solution), _p(rhs), _p(screen), _p(out),
                                 solution.shape[0], float(point_weight))
    return out, norm


def _neighbour_sum_np(x: np.ndarray):
    """capi.mojo:282-306 over a whole grid in x's dtype: terms x-1, x+1, y-1, y+1, z-1, z+1
    (grids are [z, y, x]); an absent neighbour skips its addition."""
    n = x.shape[0]
    idx = np.arange(n)
    s = np.zeros_like(x)
    deg = np.zeros(x.shape, np.int64)
    for axis in (2, 1, 0):
        for shift in (1, -1):  # roll by +1 brings index i-1 to i (the "-1" neighbour)
            v = np.roll(x, shift, axis=axis)
            ok = idx > 0 if shift == 1 else idx < n - 1
            ok = ok.reshape([n if a == axis else 1 for a in range(3)])
            ok = np.broadcast_to(ok, x.shape)
            s = np.where(ok, s + v, s)
            deg += ok
    return s, deg


def relax_np(solution, rhs, screen, point_weight: float, iterations: int) -> np.ndarray:
    """capi.mojo:309-410 in the arrays' own dtype (float32 restates the mixed-precision
    smoother): red-black sweeps, colour 0 first, x = (rhs + h^-2 sum) / diag where
    diag = deg h^-2 + (alpha screen) h^-2 > 1e-30. Every colour-c node reads only colour
    1-c nodes, so a whole-grid update per colour is the serial sweep's result."""
    x = np.array(solution, copy=True)
    dt = x.dtype.type
    n = x.shape[0]
    h2 = dt(float(n - 1) * float(n - 1))
    pw = dt(point_weight)
    zz, yy, xx = np.indices(x.shape)
    colour = (xx + yy + zz) & 1
    screening = pw * screen * h2
    for _ in range(iterations):
        for c in (0, 1):
            s, deg = _neighbour_sum_np(x)
            diag = deg.astype(dt) * h2 + screening
            with np.errstate(all="ignore"):
                new = (rhs + s * h2) / diag
            m = (colour == c) & (diag > dt(1.0e-30))
            x[m] = new[m]
    return x


def residual_np(solution, rhs, screen, point_weight: float) -> np.ndarray:
    """capi.mojo:413-470 in the arrays' own dtype: r = rhs - ((deg h^-2 + alpha s h^-2) x
    - h^-2 sum)."""
    x = np.asarray(solution)
    dt = x.dtype.type
    n = x.shape[0]
    h2 = dt(float(n - 1) * float(n - 1))
    pw = dt(point_weight)
    s, deg = _neighbour_sum_np(x)
    applied = (deg.astype(dt) * h2 + pw * screen * h2) * x - s * h2
    return rhs - applied


def restrict(fine) -> np.ndarray:
    fine = _f(fine)
    cn = (fine.shape[0] + 1) // 2
    coarse = np.empty((cn, cn, cn))
    lib().oracle_restrict(_p(fine), _p(coarse), fine.shape[0])
    return coarse


def prolong_add(coarse, fine) -> np.ndarray:
    coarse = _f(coarse)
    out = np.array(fine, dtype=np.float64, copy=True, order="C")
    lib().oracle_prolong_add(_p(coarse), _p(out), out.shape[0])
    return out


def restrict_np(fine, scale: float = 1.0) -> np.ndarray:
    """capi.mojo:556-673 in the array's own dtype: coarse (X,Y,Z) = sum of in-grid fine
    values at (2X+ox, 2Y+oy, 2Z+oz) times wx wy wz (2 at the centre, 1 aside), terms in
    (oz, oy, ox) order, divided by the sum of those weights, times `scale`."""
    fine = np.asarray(fine)
    dt = fine.dtype.type
    nf = fine.shape[0]
    nc = (nf + 1) // 2
    Z, Y, X = np.indices((nc, nc, nc))
    total = np.zeros((nc, nc, nc), dtype=fine.dtype)
    wsum = np.zeros((nc, nc, nc), dtype=fine.dtype)
    for oz in (-1, 0, 1):
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                zz, yy, xx = 2 * Z + oz, 2 * Y + oy, 2 * X + ox
                ok = (zz >= 0) & (zz < nf) & (yy >= 0) & (yy < nf) & (xx >= 0) & (xx < nf)
                v = fine[np.clip(zz, 0, nf - 1), np.clip(yy, 0, nf - 1), np.clip(xx, 0, nf - 1)]
                w = dt(2 if ox == 0 else 1) * dt(2 if oy == 0 else 1) * dt(2 if oz == 0 else 1)
                total = np.where(ok, total + v * w, total)
                wsum = np.where(ok, wsum + w, wsum)
    return total / wsum * dt(scale)


def prolong_np(coarse, fine) -> np.ndarray:
    """capi.mojo:676-742 in the arrays' own dtype: fine(x,y,z) += trilinear value of the
    coarse grid at (x/2, y/2, z/2), lerped along x, then y, then z, upper index clamped."""
    coarse = np.asarray(coarse)
    out = np.array(fine, copy=True)
    dt = out.dtype.type
    n = out.shape[0]
    nc = coarse.shape[0]
    z, y, x = np.indices(out.shape)
    x0, y0, z0 = x >> 1, y >> 1, z >> 1
    x1, y1, z1 = np.minimum(x0 + 1, nc - 1), np.minimum(y0 + 1, nc - 1), np.minimum(z0 + 1, nc - 1)
    half, zero, one = dt(0.5), dt(0.0), dt(1.0)
    tx, ty, tz = (np.where(v & 1, half, zero) for v in (x, y, z))
    c = lambda zi, yi, xi: coarse[zi, yi, xi]  # noqa: E731
    c00 = c(z0, y0, x0) * (one - tx) + c(z0, y0, x1) * tx
    c10 = c(z0, y1, x0) * (one - tx) + c(z0, y1, x1) * tx
    c01 = c(z1, y0, x0) * (one - tx) + c(z1, y0, x1) * tx
    c11 = c(z1, y1, x0) * (one - tx) + c(z1, y1, x1) * tx
    c0 = c00 * (one - ty) + c10 * ty
    c1 = c01 * (one - ty) + c11 * ty
    return out + (c0 * (one - tz) + c1 * tz)


def iso_value(field, points, normals, confidence: bool) -> float:
    field, points, normals = map(_f, (field, points, normals))
    return lib().oracle_iso_value(_p(field), _p(points), _p(normals), len(points),
                                  field.shape[0], int(confidence))


def dual_vertex_count(field, iso: float) -> int:
    field = _f(field)
    n = field.shape[0]
    scratch = np.empty(max(1, ((n - 1) + 3) // 4), dtype=np.int64)
    dummy = np.zeros(1)
    return lib().oracle_dual_vertices(_p(field), _p(field), n, float(iso), _p(scratch),
                                      _p(dummy), _p(dummy), 0)


def dual_vertices(field, density, iso: float, capacity: int):
    field, density = _f(field), _f(density)
    n = field.shape[0]
    cm = np.empty((n - 1) ** 3, dtype=np.int64)
    v = np.zeros((max(capacity, 1), 3))
    d = np.zeros(max(capacity, 1))
    cnt = lib().oracle_dual_vertices(_p(field), _p(density), n, float(iso), _p(cm), _p(v),
                                     _p(d), int(capacity))
    return cnt, cm, v[:capacity], d[:capacity]


def dual_faces(field, iso: float, cell_map, capacity: int):
    field = _f(field)
    cell_map = np.ascontiguousarray(cell_map, dtype=np.int64)
    f = np.zeros((max(capacity, 1), 3), dtype=np.int64)
    cnt = lib().oracle_dual_faces(_p(field), field.shape[0], float(iso), _p(cell_map), _p(f),
                                  int(capacity))
    return cnt, f[:capacity]


# ---- core.py restatements ------------------------------------------------------------

def valid_oriented_rows(points, normals) -> np.ndarray:
    """core.py:73-83"""
    points, normals = _f(points), _f(normals)
    finite = np.isfinite(points).all(axis=1) & np.isfinite(normals).all(axis=1)
    scale = np.max(np.abs(normals), axis=1) if len(normals) else np.zeros(0)
    rows = finite & (scale > 0)
    s2 = np.zeros(len(normals))
    q = normals[rows] / scale[rows, None]
    s2[rows] = np.einsum("ij,ij->i", q, q)
    return rows & np.isfinite(s2) & (s2 > 0)


def build_octree(points, depth: int):
    """core.py:86-127: per-level np.unique, breadth-first concatenation, parent by dict."""
    points = _f(points).reshape(-1, 3)
    if not len(points):
        z = np.array([0], np.int64)
        return dict(depth=z, morton=z.copy(), parent=np.array([-1], np.int64),
                    child_start=np.array([-1], np.int64), child_count=z.copy(),
                    point_leaf=np.empty(0, np.int64))
    codes = octree_codes(points, depth)
    levels, inverse = [], None
    for d in range(depth + 1):
        u, inv = np.unique(codes[:, d], return_inverse=True)
        levels.append(u)
        if d == depth:
            inverse = inv.astype(np.int64)
    offsets = np.cumsum([0] + [len(v) for v in levels], dtype=np.int64)
    morton = np.concatenate(levels)
    depths = np.concatenate([np.full(len(v), d, np.int64) for d, v in enumerate(levels)])
    parent = np.full(len(morton), -1, np.int64)
    child_start = np.full(len(morton), -1, np.int64)
    child_count = np.zeros(len(morton), np.int64)
    for d in range(1, depth + 1):
        prev = {int(c): int(offsets[d - 1] + i) for i, c in enumerate(levels[d - 1])}
        for local, c in enumerate(levels[d]):
            node = int(offsets[d] + local)
            p = prev[int(c) >> 3]
            parent[node] = p
            if child_start[p] < 0:
                child_start[p] = node
            child_count[p] += 1
    return dict(depth=depths, morton=morton, parent=parent, child_start=child_start,
                child_count=child_count, point_leaf=inverse + offsets[-2])


def solve(points, normals, depth: int, point_weight: float = 4.0, confidence: bool = False,
          cycles: int = 4, samples_per_node: float = 1.5, handle=None) -> dict:
    """core.py:222-306 on already-validated unit-cube rows (octree omitted). `handle`: an
    oracle library other than the default (build_mutation)."""
    points, normals = _f(points).reshape(-1, 3), _f(normals).reshape(-1, 3)
    valid = valid_oriented_rows(points, normals)
    points, normals = np.ascontiguousarray(points[valid]), np.ascontiguousarray(normals[valid])
    n = (1 << depth) + 1
    field, density, screen, rhs = (np.empty((n, n, n)) for _ in range(4))
    iso, rn = ctypes.c_double(), ctypes.c_double()
    acc = (handle or lib()).oracle_solve(
        _p(points), _p(normals), len(points), depth, float(point_weight), int(confidence),
        int(cycles), float(samples_per_node), _p(field), _p(density), _p(screen), _p(rhs),
        ctypes.addressof(iso), ctypes.addressof(rn))
    return dict(field=field, density=density, screen=screen, rhs=rhs, iso_value=iso.value,
                residual_norm=rn.value, accepted=acc, points=points, normals=normals)


def extract_isosurface(field, iso_value: float = 0.0, density=None):
    """core.py:309-374 count-then-fill protocol; returns (vertices, faces, vertex_density)."""
    field = _f(field)
    density = np.zeros_like(field) if density is None else _f(density)
    count = dual_vertex_count(field, iso_value)
    if count == 0:
        return np.empty((0, 3)), np.empty((0, 3), np.int64), np.empty(0)
    written, cm, v, d = dual_vertices(field, density, iso_value, count)
    assert written == count
    fcount, _ = dual_faces(field, iso_value, cm, 0)
    fwritten, f = dual_faces(field, iso_value, cm, fcount)
    assert fwritten == fcount
    return v, f.reshape(-1, 3), d


def normalize(points, scale: float):
    """core.py:377-385"""
    mn, mx = points.min(axis=0), points.max(axis=0)
    center = 0.5 * (mn + mx)
    width = float(np.max(mx - mn)) * scale
    if not np.isfinite(width) or width <= 0:
        width = 1.0
    origin = center - 0.5 * width
    return np.ascontiguousarray((points - origin) / width), origin, width


def reconstruct(points, normals, depth: int = 6, point_weight: float = 4.0,
                confidence: bool = False, cycles: int = 4, scale: float = 1.1,
                samples_per_node: float = 1.5, handle=None) -> dict:
    """core.py:388-426 (`handle`: the library whose solve runs, e.g. openmp())"""
    points, normals = _f(points).reshape(-1, 3), _f(normals).reshape(-1, 3)
    valid = valid_oriented_rows(points, normals)
    points, normals = points[valid], normals[valid]
    if not len(points):
        return dict(vertices=np.empty((0, 3)), faces=np.empty((0, 3), np.int64),
                    density=np.empty(0), iso_value=0.0, residual_norm=0.0)
    unit, origin, width = normalize(points, scale)
    sol = solve(unit, normals, depth, point_weight, confidence, cycles, samples_per_node,
                handle=handle)
    v, f, d = extract_isosurface(sol["field"], sol["iso_value"], sol["density"])
    return dict(vertices=np.ascontiguousarray(v * width + origin), faces=f, density=d,
                iso_value=sol["iso_value"], residual_norm=sol["residual_norm"], solution=sol,
                origin=origin, width=width)


def oriented_points_from_mesh_c(vertices, faces):
    """core.py:429-450 as the C per-face loop (mpr_oracle.c), for large meshes."""
    vertices = _f(vertices).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int64).reshape(-1, 3)
    p, n = np.empty_like(vertices), np.empty_like(vertices)
    k = lib().oracle_oriented_points(_p(vertices), len(vertices), _p(faces), len(faces),
                                     _p(p), _p(n))
    return np.ascontiguousarray(p[:k]), np.ascontiguousarray(n[:k])


def oriented_points_from_mesh(vertices, faces):
    """core.py:429-450 (per-face loop, unnormalised cross products)."""
    vertices = _f(vertices).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int64).reshape(-1, 3)
    normals = np.zeros_like(vertices)
    referenced = np.zeros(len(vertices), dtype=bool)
    for face in faces:
        a, b, c = map(int, face)
        cross = np.cross(vertices[b] - vertices[a], vertices[c] - vertices[a])
        if np.isfinite(cross).all() and np.dot(cross, cross) > 0:
            normals[a] += cross
            normals[b] += cross
            normals[c] += cross
            referenced[[a, b, c]] = True
    lengths = np.linalg.norm(normals, axis=1)
    keep = referenced & (lengths > 0) & np.isfinite(lengths)
    normals[keep] /= lengths[keep, None]
    return np.ascontiguousarray(vertices[keep]), np.ascontiguousarray(normals[keep])
