"""CPU tests of the served request stream's definition (the request mix the server is
measured on): Poisson arrivals at the offered rate, depths 6:7:8:9 weighted 4:3:2:1, clouds
of 5k-500k points rotated and scaled into [-1, 1]^3, deterministic per seed, and the
schedule/cloud split the bench's client threads use giving the same requests."""

from __future__ import annotations

import numpy as np

from mojo_poissonrecon import shapes


def test_served_stream_matches_its_definition():
    count, qps = 4000, 250.0
    base, items = shapes.served_schedule(count, qps, seed=11, pool=6)
    assert len(base) == 6 and len(items) == count
    arrivals = np.array([it[0] for it in items])
    assert np.all(np.diff(arrivals) > 0)
    rate = count / arrivals[-1]
    assert abs(rate - qps) / qps < 0.08  # Poisson: 4000 arrivals, about 1.6 % rel. sd
    depths = np.array([it[4] for it in items])
    assert set(depths.tolist()) <= {6, 7, 8, 9}
    share = np.array([(depths == d).mean() for d in (6, 7, 8, 9)])
    assert np.allclose(share, [0.4, 0.3, 0.2, 0.1], atol=0.03)
    for p, n in base:
        assert 5000 <= len(p) <= 500_000 and p.shape == n.shape
    for it in items[:20]:
        arrival, p, n, d = shapes.served_cloud(base, it)
        rot = it[2]
        assert np.allclose(rot @ rot.T, np.eye(3), atol=1e-12) and np.linalg.det(rot) > 0
        assert np.abs(p).max() <= 1.0 + 1e-12 and p.dtype == np.float64
        assert p.flags.c_contiguous and n.flags.c_contiguous
        assert arrival == it[0] and d == it[4]


def test_served_requests_are_the_schedule_and_are_deterministic():
    a = list(shapes.served_requests(30, 100.0, seed=4, pool=3))
    base, items = shapes.served_schedule(30, 100.0, seed=4, pool=3)
    b = [shapes.served_cloud(base, it) for it in items]
    c = list(shapes.served_requests(30, 100.0, seed=4, pool=3))
    for x, y, z in zip(a, b, c):
        assert x[0] == y[0] == z[0] and x[3] == y[3] == z[3]
        assert np.array_equal(x[1], y[1]) and np.array_equal(x[1], z[1])
        assert np.array_equal(x[2], y[2]) and np.array_equal(x[2], z[2])
    d = list(shapes.served_requests(30, 100.0, seed=5, pool=3))
    assert any(x[0] != y[0] for x, y in zip(a, d))
