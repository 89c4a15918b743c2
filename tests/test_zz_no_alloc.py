"""GPU test of the allocation-free iso value (mpr_iso_value_f64_noalloc).

The entry point takes device buffers and keeps its partial sums in the library's static
device array. It must give the value of mpr_iso_value_f64 (the same kernels and block
layout, so bit for bit) on two calls in a row, the second smaller than the first, and reject
a point count beyond its static scratch. Device buffers come from the HIP runtime the
library itself is linked against.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from mojo_poissonrecon._lib import addr, last_error, lib

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _DeviceBuffers:
    """hipMalloc'd copies of host arrays, freed on exit."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def upload(self, array: np.ndarray) -> int:
        array = np.ascontiguousarray(array)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(8, array.nbytes)) == 0
        self.ptrs.append(p.value)
        if array.nbytes:
            assert self.hip.hipMemcpy(p.value, array.ctypes.data, array.nbytes, _H2D) == 0
        return p.value

    def download_double(self, device: int) -> float:
        out = np.zeros(1)
        assert self.hip.hipMemcpy(out.ctypes.data, device, 8, _D2H) == 0
        return float(out[0])

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _noalloc_fn():
    return lib().mpr_iso_value_f64_noalloc


def _case(seed: int, n: int, count: int):
    rng = np.random.default_rng(seed)
    field = rng.normal(size=(n, n, n))
    points = rng.random((count, 3))
    points[:4] = [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.0, 0.5], [0.5, 1.0, 0.0]]
    normals = rng.normal(size=(count, 3)) * rng.uniform(0.5, 1.5, (count, 1))
    return field, points, normals


def test_noalloc_iso_equals_the_allocating_entry_point_on_two_calls():
    fn = _noalloc_fn()
    with _DeviceBuffers() as dev:
        out = dev.upload(np.full(1, np.nan))
        # the second call is smaller (fewer blocks, smaller grid) and must not see the first
        # call's partial sums, which stay in the static array
        for seed, n, count, confidence in ((1, 65, 300_000, 1), (2, 33, 3_000, 0)):
            field, points, normals = _case(seed, n, count)
            want = lib().mpr_iso_value_f64(addr(field), addr(points), addr(normals), count, n,
                                           confidence)
            status = fn(dev.upload(field), dev.upload(points), dev.upload(normals), count, n,
                        confidence, out, None)
            assert status == 0, last_error()
            assert dev.hip.hipDeviceSynchronize() == 0
            got = dev.download_double(out)
            assert got == want  # same kernels, same partial-sum layout
            expected = oracle.iso_value(field, points, normals, bool(confidence))
            assert got == pytest.approx(expected, rel=1e-12)


def test_noalloc_iso_of_no_points_is_zero_and_oversize_is_rejected():
    fn = _noalloc_fn()
    with _DeviceBuffers() as dev:
        field = np.ones((9, 9, 9))
        out = dev.upload(np.full(1, np.nan))
        dfield = dev.upload(field)
        assert fn(dfield, None, None, 0, 9, 0, out, None) == 0, last_error()
        assert dev.hip.hipDeviceSynchronize() == 0
        assert dev.download_double(out) == 0.0
        # beyond the static scratch (2^24 points): an error, nothing launched
        assert fn(dfield, dfield, dfield, (1 << 24) + 1, 9, 0, out, None) == -1
        assert "16777216" in last_error()
        assert fn(dfield, dfield, dfield, 5, 1, 0, out, None) == -1
        assert dev.download_double(out) == 0.0
