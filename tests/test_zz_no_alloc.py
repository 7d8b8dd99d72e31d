"""The allocation-free entry point icempm_sort_raw_points_noalloc against
icempm_sort_raw_points, through the C ABI: the same stable cell-key order
(bit-exact, integer and byte work) on two calls in a row, the second with a
smaller input, so the shared static scratch keeps the first call's contents
when the second runs; and a size past ICEMPM_NOALLOC_MAX_POINTS refused on
the host with a message.  Also checked against the oracle's stable order."""
import ctypes
from ctypes import POINTER, c_float, c_int16

import numpy as np
import pytest

import oracle
from icempm_amd import _lib
from icempm_amd._lib import ICEMPM_E_ARG

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1048576   # ICEMPM_NOALLOC_MAX_POINTS in include/icempm.h


def raw_cloud(n, seed, gridx, h):
    rng = np.random.default_rng(seed)
    xy = np.empty((n, 2), dtype=np.float32)
    xy[:, 0] = rng.uniform(0.5 * h, (gridx - 3) * h, n)
    xy[:, 1] = rng.uniform(0.5 * h, 60 * h, n)
    g = rng.integers(-32768, 32767, n).astype(np.int16)
    return xy, g


def sort_with(fn, ctx, xy, g, hinv, gridx):
    a = np.ascontiguousarray(xy.copy())
    b = np.ascontiguousarray(g.copy())
    rc = fn(ctx, a.ctypes.data_as(POINTER(c_float)), b.ctypes.data_as(POINTER(c_int16)), a.shape[0],
            hinv, gridx)
    return rc, a, b


def test_noalloc_sort_equals_allocating_sort(gpu_available):
    H = _lib.hip()
    ctx = ctypes.c_void_p()
    assert H.icempm_create(ctypes.byref(ctx)) == 0
    try:
        h, gridx = 0.005, 1200
        hinv = 1.0 / h
        for n, seed in ((700_001, 3), (65_772, 4)):     # the second call smaller
            print("noalloc sort: n", n, "seed", seed, flush=True)
            xy, g = raw_cloud(n, seed, gridx, h)
            rc1, xa, ga = sort_with(H.icempm_sort_raw_points, ctx, xy, g, hinv, gridx)
            rc2, xb, gb = sort_with(H.icempm_sort_raw_points_noalloc, ctx, xy, g, hinv, gridx)
            assert rc1 == 0 and rc2 == 0, (rc1, rc2, H.icempm_last_error(ctx))
            assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
            assert np.array_equal(ga, gb)
            perm = oracle.stable_order(oracle.cell_keys_f32(xy, hinv, gridx))
            assert np.array_equal(xb.view(np.uint32), xy[perm].view(np.uint32))
            assert np.array_equal(gb, g[perm])
        print("noalloc sort: size past the maximum", flush=True)
        xy, g = raw_cloud(16, 5, gridx, h)
        big = np.zeros((NOALLOC_MAX + 1, 2), dtype=np.float32)
        bg = np.zeros(NOALLOC_MAX + 1, dtype=np.int16)
        rc = H.icempm_sort_raw_points_noalloc(ctx, big.ctypes.data_as(POINTER(c_float)),
                                              bg.ctypes.data_as(POINTER(c_int16)), NOALLOC_MAX + 1, hinv, gridx)
        assert rc == ICEMPM_E_ARG
        assert b"ICEMPM_NOALLOC_MAX_POINTS" in H.icempm_last_error(ctx)
        # the largest size it takes
        xy, g = raw_cloud(NOALLOC_MAX, 6, gridx, h)
        rc1, xa, ga = sort_with(H.icempm_sort_raw_points, ctx, xy, g, hinv, gridx)
        rc2, xb, gb = sort_with(H.icempm_sort_raw_points_noalloc, ctx, xy, g, hinv, gridx)
        assert rc1 == 0 and rc2 == 0
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)) and np.array_equal(ga, gb)
    finally:
        H.icempm_destroy(ctx)
