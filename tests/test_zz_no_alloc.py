"""The no-allocation entry point (bfj_self_join_device_noalloc) against the allocating
one (bfj_self_join_device), through the C ABI.

Both run the same kernels; the no-allocation call takes every working buffer from a
file-scope device array instead of hipMalloc.  Per-point counts and totals must be
identical (they are exact), on two calls in a row, the second on a smaller dataset
(the array keeps the first call's contents), for every dtype; a dataset beyond the
array's size is refused with BFJ_ERR_CAPACITY and leaves the context usable, and a
candidate stream beyond the array's share runs in slices within it.
"""
import ctypes

import numpy as np
import pytest

import tensor_bf_main_amd as bfj
from oracle import oracle
from tensor_bf_main_amd import workloads as wl

pytestmark = pytest.mark.gpu
DTYPES = ["fp64", "fp16", "bf16", "fp8"]


def _call(ctx, fn, xd, n, d, eps, dtype, rows=None):
    e = np.ascontiguousarray(eps, dtype=np.float64)
    rb, re = rows if rows else (0, n)
    cnt = ctx.empty((e.size, re - rb), np.int64)
    tot = np.zeros(e.size, dtype=np.uint64)
    st = bfj.Stats()
    rc = fn(ctx._h, xd.data_ptr(), n, d, d, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), e.size,
            bfj.DTYPES[dtype], rb, re, cnt.data_ptr(), tot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            ctypes.byref(st))
    out = (rc, cnt.numpy() if rc == 0 else None, tot, st.as_dict())
    cnt.free()
    return out


@pytest.mark.parametrize("dtype", DTYPES)
def test_noalloc_matches_the_allocating_entry_point(ctx, dtype):
    lib = ctx.lib
    # (The e4m3 band of FP8 passes is large: quantiles kept small enough that every
    # case fits the 64 MiB scratch for every dtype.)
    cases = [(wl.gmm(4000, 64, 8, 0.05, 3), None, [0.0005, 0.003]),   # first call
             (wl.uniform(2500, 40, 4), None, [0.002, 0.01]),           # second call, smaller
             (wl.uniform(3000, 128, 5), (700, 2300), [0.001, 0.01])]   # a row range (shard)
    for x, rows, qs in cases:
        n, d = x.shape
        eps = wl.eps_quantile(x, qs, 300, n)
        xd = ctx.to_device(x)
        rc0, c0, t0, s0 = _call(ctx, lib.bfj_self_join_device_noalloc, xd, n, d, eps, dtype, rows)
        assert rc0 == 0, lib.bfj_last_error()
        rc1, c1, t1, s1 = _call(ctx, lib.bfj_self_join_device, xd, n, d, eps, dtype, rows)
        assert rc1 == 0, lib.bfj_last_error()
        np.testing.assert_array_equal(c0, c1)
        np.testing.assert_array_equal(t0, t1)
        assert s0["band_pairs"] == s1["band_pairs"] or dtype != "fp64"
        rb, re = rows if rows else (0, n)
        if n <= 3000:
            oc, ot = oracle.join(x, eps, rb, re)
            np.testing.assert_array_equal(c0, oc)
            np.testing.assert_array_equal(t0, ot)
        xd.free()


def test_noalloc_refuses_what_does_not_fit(ctx):
    lib = ctx.lib
    x = wl.uniform(200_000, 128, 6)
    xd = ctx.to_device(x)
    rc, _, _, _ = _call(ctx, lib.bfj_self_join_device_noalloc, xd, 200_000, 128, [0.5], "fp16")
    assert rc == bfj.BFJ_ERR_CAPACITY
    assert b"no-allocation scratch" in lib.bfj_last_error()
    xd.free()
    # The context and the no-allocation path still work afterwards.
    y = wl.uniform(1500, 16, 7)
    yd = ctx.to_device(y)
    eps = wl.eps_quantile(y, [0.01], 200, 1)
    rc, c, t, _ = _call(ctx, lib.bfj_self_join_device_noalloc, yd, 1500, 16, eps, "fp16")
    assert rc == 0
    oc, ot = oracle.join(y, eps)
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(t, ot)
    yd.free()
    res = ctx.self_join(y, eps, dtype="fp16")
    np.testing.assert_array_equal(res.counts, oc)


@pytest.mark.parametrize("dtype", ["fp16", "fp64"])
def test_noalloc_slices_a_stream_beyond_its_array(ctx, dtype):
    """eps at the 40 % distance quantile: about 3.2e6 candidate records, more than the
    array's share for the stream, so the call runs the join in slices within it.  The
    counts equal the allocating call's and the oracle's."""
    lib = ctx.lib
    x = wl.uniform(4000, 64, 12)
    eps = wl.eps_quantile(x, [0.05, 0.4], 300, 12)
    xd = ctx.to_device(x)
    rc0, c0, t0, s0 = _call(ctx, lib.bfj_self_join_device_noalloc, xd, 4000, 64, eps, dtype)
    assert rc0 == 0, lib.bfj_last_error()
    assert s0["slices"] >= 2, s0
    rc1, c1, t1, s1 = _call(ctx, lib.bfj_self_join_device, xd, 4000, 64, eps, dtype)
    assert rc1 == 0, lib.bfj_last_error()
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(t0, t1)
    assert s0["candidates"] == s1["candidates"] and s0["band_pairs"] == s1["band_pairs"]
    oc, ot = oracle.join(x, eps)
    np.testing.assert_array_equal(c0, oc)
    np.testing.assert_array_equal(t0, ot)
    xd.free()
