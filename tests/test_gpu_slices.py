"""Passes whose candidate stream exceeds the stream limit run the join in slices.

bfj_set_stream_limit caps the device memory of the candidate stream and band lists;
a pass that needs more runs the join's work items in slices, each followed by its band
tests, middle stage and FP64 refine.  Every pair is decided by exactly one record of
one slice, so the per-point counts, the totals and the pass statistics (candidates,
band pairs, pairs left for FP64) must equal those of the same pass in one piece, and
the oracle's.  Limits here are small enough to force several slices at small n; each
test uses its own context so that the shrunken capacities stay with it.
"""
import numpy as np
import pytest

import tensor_bf_main_amd as bfj
from oracle import oracle
from tensor_bf_main_amd import workloads as wl

pytestmark = pytest.mark.gpu
DTYPES = ["fp64", "fp16", "bf16", "fp8"]

C = bfj.CHUNK_RECORDS  # candidate records (16 bytes each) per chunk


def _limit_for(stats, dtype, pieces):
    """A stream limit whose slices hold about 1/pieces of the pass's candidate chunks:
    per chunk 16C bytes of records and 4/3 x C + 64 band slots of 16 B per band list,
    plus 1025 slots per list.  At pieces >= 5 the pass in one piece does not fit it."""
    lists = 1 if dtype == "fp64" else 2
    chunks = max(1, stats["candidates"] // C // pieces)
    per_chunk = 16 * C + lists * 16 * (C * 4 / 3 + 64)
    return int(lists * 16 * 1025 + chunks * per_chunk) + 1


@pytest.fixture
def own_ctx():
    c = bfj.Context(0)
    yield c
    c.close()


def _same_stats(a, b):
    for k in ("candidates", "band_pairs", "fp64_pairs", "pairs"):
        assert a[k] == b[k], (k, a[k], b[k])


@pytest.mark.parametrize("dtype", DTYPES)
def test_sliced_pass_equals_one_piece(own_ctx, dtype):
    x, eps = wl.config(3, n=9000)
    one = own_ctx.self_join(x, eps, dtype=dtype)
    assert one.stats["slices"] == 1
    own_ctx.set_stream_limit(_limit_for(one.stats, dtype, 5))
    sl = own_ctx.self_join(x, eps, dtype=dtype)
    assert sl.stats["slices"] >= 3, sl.stats
    oc, ot = oracle.join(x, eps)
    np.testing.assert_array_equal(sl.counts, oc)
    np.testing.assert_array_equal(sl.totals, ot)
    np.testing.assert_array_equal(one.counts, oc)
    _same_stats(one.stats, sl.stats)
    # The same context again: straight into slices, same result.
    again = own_ctx.self_join(x, eps, dtype=dtype)
    np.testing.assert_array_equal(again.counts, oc)
    assert again.stats["slices"] >= 3
    # Back to the default limit: one piece.
    own_ctx.set_stream_limit(0)
    back = own_ctx.self_join(x, eps, dtype=dtype)
    np.testing.assert_array_equal(back.counts, oc)
    assert back.stats["slices"] == 1


@pytest.mark.parametrize("dtype", ["fp16", "fp64"])
def test_sliced_row_range_and_wide_eps(own_ctx, dtype):
    """A row range (one shard's query rows) and eps up to the 30 % distance quantile,
    where most pairs are candidates."""
    x = wl.gmm(5000, 32, 8, 0.05, 41)
    eps = wl.eps_quantile(x, [0.01, 0.3], 300, 5)
    rb, re = 1200, 3900
    one = own_ctx.self_join(x, eps, dtype=dtype, row_begin=rb, row_end=re)
    own_ctx.set_stream_limit(_limit_for(one.stats, dtype, 7))
    sl = own_ctx.self_join(x, eps, dtype=dtype, row_begin=rb, row_end=re)
    assert sl.stats["slices"] >= 4, sl.stats
    oc, ot = oracle.join(x, eps, rb, re)
    np.testing.assert_array_equal(sl.counts, oc)
    np.testing.assert_array_equal(sl.totals, ot)
    _same_stats(one.stats, sl.stats)


@pytest.mark.parametrize("dtype", ["fp16", "fp8"])
def test_sliced_pair_list(own_ctx, dtype):
    x, eps = wl.config(3, n=6000)
    e0 = float(np.atleast_1d(eps)[0])
    p1, t1, s1 = own_ctx.self_join_pairs(x, e0, dtype=dtype)
    own_ctx.set_stream_limit(_limit_for(s1, dtype, 6))
    p2, t2, s2 = own_ctx.self_join_pairs(x, e0, dtype=dtype)
    assert s2["slices"] >= 2, s2
    assert t1 == t2
    a = np.unique(np.asarray(p1).reshape(-1, 2), axis=0)
    b = np.unique(np.asarray(p2).reshape(-1, 2), axis=0)
    np.testing.assert_array_equal(a, b)
    assert len(a) == len(np.asarray(p2).reshape(-1, 2))


@pytest.mark.parametrize("dtype", ["fp16", "fp8"])
def test_sliced_plan(own_ctx, dtype):
    """A batch plan: one that needs slices from its first pass runs them eagerly on every
    run; a captured plan whose stream later outgrows the limit switches to slices."""
    sets = wl.batch(n_sets=5, seed=91, n_min=1000, n_max=3000)
    dev = [{"data": own_ctx.to_device(s["data"]), "eps": s["eps"],
            "counts": own_ctx.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
    ref = [oracle.join(s["data"], s["eps"]) for s in sets]

    def check(plan, st):
        for dv, (oc, ot), tot in zip(dev, ref, plan.totals()):
            np.testing.assert_array_equal(dv["counts"].numpy(), oc)
            np.testing.assert_array_equal(tot, ot)
        return st

    whole = own_ctx.plan(dev, dtype=dtype)
    st0 = check(whole, whole.run())
    assert st0["slices"] == 1
    # Captured plan, then a small limit and a small stream: the replay overflows and
    # the plan goes on in slices.
    own_ctx.set_stream_limit(_limit_for(st0, dtype, 6))
    own_ctx.set_capacity(candidate_chunks=1)
    st1 = check(whole, whole.run())
    assert st1["slices"] >= 2, st1
    _same_stats(st0, st1)
    st2 = check(whole, whole.run())
    assert st2["slices"] >= 2
    whole.close()
    # A plan made under the small limit.
    sliced = own_ctx.plan(dev, dtype=dtype)
    st3 = check(sliced, sliced.run())
    assert st3["slices"] >= 2
    _same_stats(st0, st3)
    sliced.close()
    for dv in dev:
        dv["data"].free()
        dv["counts"].free()


def test_limit_too_small_is_refused(own_ctx):
    """A limit below one chunk and its band slots, then one that a single work item's
    candidates exceed: both fail with BFJ_ERR_CAPACITY and leave the context usable."""
    x, eps = wl.config(3, n=4000)
    per_chunk = 16 * C + 2 * 16 * (C * 4 / 3 + 64)
    for limit, what in ((2 * 16 * 1025 + int(0.5 * per_chunk), "stream limit"),
                        (2 * 16 * 1025 + int(2.03 * per_chunk), "one join work item")):
        own_ctx.set_stream_limit(limit)
        with pytest.raises(bfj.JoinError) as e:
            own_ctx.self_join(x, eps, dtype="fp16")
        assert e.value.code == bfj.BFJ_ERR_CAPACITY
        assert what in str(e.value), str(e.value)
    own_ctx.set_stream_limit(0)
    res = own_ctx.self_join(x, eps, dtype="fp16")
    oc, ot = oracle.join(x, eps)
    np.testing.assert_array_equal(res.counts, oc)
    np.testing.assert_array_equal(res.totals, ot)
    assert res.stats["slices"] == 1


def test_growth_that_cannot_allocate_runs_in_slices(own_ctx):
    """A pass whose candidate stream must grow, with the k-th device allocation from
    then on failing as out-of-memory: a failed growth of the stream falls back to a
    sliced pass under half the bytes that failed (exact counts), and any other failed
    allocation reports BFJ_ERR_NOMEM and leaves the context usable."""
    lib = own_ctx.lib
    x, eps = wl.config(3, n=9000)
    oc, ot = oracle.join(x, eps)
    one = own_ctx.self_join(x, eps, dtype="fp16")
    sliced = 0
    try:
        for k in range(1, 9):
            # Stream buffers given back (a limit below them), the default limit again,
            # and capacities that the pass must outgrow.
            own_ctx.set_stream_limit(1)
            own_ctx.set_stream_limit(0)
            own_ctx.set_capacity(candidate_chunks=1, band_records=8)
            assert lib.bfj_fail_allocation(k) == 0
            try:
                res = own_ctx.self_join(x, eps, dtype="fp16")
            except bfj.JoinError as err:
                assert err.code == bfj.BFJ_ERR_NOMEM and "injected" in str(err), str(err)
                continue
            finally:
                lib.bfj_fail_allocation(0)
            np.testing.assert_array_equal(res.counts, oc, err_msg=f"k = {k}")
            np.testing.assert_array_equal(res.totals, ot)
            _same_stats(one.stats, res.stats)
            sliced += res.stats["slices"] >= 2
    finally:
        lib.bfj_fail_allocation(0)
    assert sliced >= 1
    own_ctx.set_stream_limit(0)
    res = own_ctx.self_join(x, eps, dtype="fp16")
    np.testing.assert_array_equal(res.counts, oc)
