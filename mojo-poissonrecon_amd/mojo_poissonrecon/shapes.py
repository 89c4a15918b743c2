"""Seeded synthetic oriented point clouds for the benchmark and parity workloads.

* sphere: z ~ U(-1, 1), theta ~ U(0, 2 pi), normals = positions (README.md:64-71 and
  bench/bench.py:34-41 of the reference);
* torus: area-uniform samples (rejection on (R + r cos v) / (R + r)), R = 0.6, r = 0.25;
* superquadric: signed-power parametric samples, normals from the implicit gradient.

Noise: isotropic Gaussian position noise with sigma given as a fraction of the bbox
diagonal; normal noise rotates each unit normal by an angle ~ N(0, deg) about a uniformly
random axis perpendicular to it; normal lengths are drawn from ``length_range``.
"""

from __future__ import annotations

import math

import numpy as np


def sphere(count: int, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, count)
    theta = rng.uniform(0.0, 2.0 * np.pi, count)
    radius = np.sqrt(1.0 - z * z)
    p = np.ascontiguousarray(np.column_stack((radius * np.cos(theta), radius * np.sin(theta), z)))
    return p, p.copy()


def torus(count: int, seed: int = 0, R: float = 0.6, r: float = 0.25):
    rng = np.random.default_rng(seed)
    us, vs = [], []
    need = count
    while need > 0:
        u = rng.uniform(0, 2 * np.pi, 2 * need + 16)
        v = rng.uniform(0, 2 * np.pi, 2 * need + 16)
        keep = rng.uniform(0, 1, u.size) < (R + r * np.cos(v)) / (R + r)
        us.append(u[keep][:need])
        vs.append(v[keep][:need])
        need -= int(min(keep.sum(), need))
    u, v = np.concatenate(us)[:count], np.concatenate(vs)[:count]
    p = np.column_stack(((R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u),
                         r * np.sin(v)))
    nrm = np.column_stack((np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)))
    return np.ascontiguousarray(p), np.ascontiguousarray(nrm)


def _spow(x, e):
    return np.sign(x) * np.abs(x) ** e


def superquadric(count: int, seed: int = 0, e1: float = 0.3, e2: float = 0.3,
                 half=(0.8, 0.6, 0.4)):
    rng = np.random.default_rng(seed)
    eta = rng.uniform(-np.pi / 2, np.pi / 2, count)
    omega = rng.uniform(-np.pi, np.pi, count)
    a, b, c = half
    ce, se, cw, sw = np.cos(eta), np.sin(eta), np.cos(omega), np.sin(omega)
    p = np.column_stack((a * _spow(ce, e1) * _spow(cw, e2), b * _spow(ce, e1) * _spow(sw, e2),
                         c * _spow(se, e1)))
    # gradient of F = ((x/a)^(2/e2) + (y/b)^(2/e2))^(e2/e1) + (z/c)^(2/e1)
    nrm = np.column_stack((_spow(ce, 2 - e1) * _spow(cw, 2 - e2) / a,
                           _spow(ce, 2 - e1) * _spow(sw, 2 - e2) / b,
                           _spow(se, 2 - e1) / c))
    length = np.linalg.norm(nrm, axis=1)
    bad = ~(length > 0)
    nrm[bad] = p[bad]
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    return np.ascontiguousarray(p), np.ascontiguousarray(nrm)


def perturb(points, normals, seed: int = 0, sigma: float = 0.0, degrees: float = 0.0,
            length_range=(1.0, 1.0)):
    """Position noise (sigma x bbox diagonal), angular normal noise and normal lengths."""
    rng = np.random.default_rng(seed + 7919)
    p = np.array(points, dtype=np.float64)
    n = np.array(normals, dtype=np.float64)
    n /= np.linalg.norm(n, axis=1)[:, None]
    if sigma > 0:
        diag = float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))
        p += rng.normal(0.0, sigma * diag, p.shape)
    if degrees > 0:
        t = rng.normal(size=n.shape)
        t -= np.einsum("ij,ij->i", t, n)[:, None] * n
        t /= np.linalg.norm(t, axis=1)[:, None]
        ang = np.deg2rad(rng.normal(0.0, degrees, len(n)))
        n = n * np.cos(ang)[:, None] + t * np.sin(ang)[:, None]
        n /= np.linalg.norm(n, axis=1)[:, None]
    lo, hi = length_range
    if lo != 1.0 or hi != 1.0:
        n *= rng.uniform(lo, hi, len(n))[:, None]
    return np.ascontiguousarray(p), np.ascontiguousarray(n)


def torus_workload(count: int = 200_000, seed: int = 0):
    """Single FP64 depth-8 workload: 200k torus points, sigma 0.002 diag, 5 deg normal noise,
    normal lengths U(0.5, 1.5), used with confidence weighting."""
    p, n = torus(count, seed)
    return perturb(p, n, seed, sigma=0.002, degrees=5.0, length_range=(0.5, 1.5))


def superquadric_workload(count: int = 2_000_000, seed: int = 0):
    """Mixed-precision depth-9 workload: 2M superquadric points, sigma 0.001, 3 deg."""
    p, n = superquadric(count, seed)
    return perturb(p, n, seed, sigma=0.001, degrees=3.0)


def uneven_batch_workload(count: int = 64, seed: int = 1):
    """Batch of independent clouds: depths U{5..8}, points log-uniform [2k, 200k], shapes
    cycling sphere -> torus -> superquadric, position noise U[0, 0.003] x diagonal. Returns
    (clouds, depths)."""
    rng = np.random.default_rng(seed)
    makers = (sphere, torus, superquadric)
    clouds, depths = [], []
    for i in range(count):
        npts = int(round(math.exp(rng.uniform(math.log(2e3), math.log(2e5)))))
        p, n = makers[i % 3](npts, seed=seed * 1000 + i)
        p, n = perturb(p, n, seed=seed * 1000 + i, sigma=float(rng.uniform(0, 0.003)))
        clouds.append((p, n))
        depths.append(int(rng.integers(5, 9)))
    return clouds, depths


def served_schedule(count: int, qps: float, seed: int = 2, pool: int = 24):
    """The served stream of served_requests without its point arrays: the `pool` base clouds
    and, per request, (arrival, base index, rotation, scale, depth). served_cloud(base, item)
    forms a request's rows, so a client can form them ahead of time on other threads."""
    rng = np.random.default_rng(seed)
    makers = (sphere, torus, superquadric)
    base = []
    for i in range(pool):
        npts = int(round(math.exp(rng.uniform(math.log(5e3), math.log(5e5)))))
        p, n = makers[i % 3](npts, seed=7000 + i)
        base.append(perturb(p, n, seed=7000 + i, sigma=0.001, degrees=2.0))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, count))
    depths = rng.choice([6, 7, 8, 9], size=count, p=[0.4, 0.3, 0.2, 0.1])
    picks = rng.integers(0, pool, count)
    items = []
    for k in range(count):
        rot = random_rotation(rng)
        scale = float(rng.uniform(0.5, 1.0))
        items.append((float(arrivals[k]), int(picks[k]), rot, scale, int(depths[k])))
    return base, items


def served_cloud(base, item):
    """(arrival, points, normals, depth) of one served_schedule item."""
    arrival, pick, rot, scale, depth = item
    p, n = base[pick]
    pts = p @ rot.T
    pts = pts / np.abs(pts).max() * scale
    return arrival, np.ascontiguousarray(pts), np.ascontiguousarray(n @ rot.T), depth


def served_requests(count: int, qps: float, seed: int = 2, pool: int = 24):
    """A served request stream: Poisson arrivals at `qps` (arrival times in seconds from 0),
    each a cloud of 5k-500k points (log-uniform) on a random shape family, randomly rotated
    and scaled into [-1, 1]^3, at a depth from {6, 7, 8, 9} weighted 4:3:2:1. The clouds are
    drawn from `pool` pre-generated ones (rotated and scaled anew per request) to bound the
    host work; yields (arrival, points, normals, depth) lazily."""
    base, items = served_schedule(count, qps, seed, pool)
    for item in items:
        yield served_cloud(base, item)


def random_rotation(rng) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q
