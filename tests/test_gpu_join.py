"""GPU parity tests: the HIP join (through the C ABI) against the CPU FP64 oracle.

The bar is bit-exact per-point counts and totals for every dtype: the
low-precision paths decide only the pairs whose rigorous error band excludes
the threshold and re-evaluate the rest in FP64 on the GPU.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import tensor_bf_main_amd as bfj
from oracle import oracle
from tensor_bf_main_amd import workloads as wl

pytestmark = pytest.mark.gpu
DTYPES = ["fp64", "fp16", "bf16", "fp8"]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _check(ctx, x, eps, dtype, row_begin=0, row_end=None):
    n = x.shape[0]
    re_ = n if row_end is None else row_end
    res = ctx.self_join(x, eps, dtype=dtype, row_begin=row_begin, row_end=re_)
    oc, ot = oracle.join(x, eps, row_begin, re_)
    np.testing.assert_array_equal(res.counts, oc, err_msg=f"per-point counts differ ({dtype})")
    np.testing.assert_array_equal(res.totals, ot, err_msg=f"totals differ ({dtype})")
    return res


@pytest.mark.parametrize("dtype", DTYPES)
def test_reference_test_files(ctx, dtype):
    gold = json.load(open(os.path.join(GOLDEN, "expected_counts.json")))
    for name in ("test100.txt", "test128.txt"):
        rows = oracle.parse_text(open(os.path.join(GOLDEN, name)).read(), 18)
        g = gold[name]
        res = ctx.self_join(rows, g["eps"], dtype=dtype)
        assert [int(t) for t in res.totals] == g["totals"]
        np.testing.assert_array_equal(res.counts, np.asarray(g["counts"], dtype=np.int64))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,d", [(1, 1), (2, 3), (17, 5), (300, 18), (1000, 32), (2000, 128), (1500, 200),
                                 (1200, 256)])
def test_uniform_small(ctx, dtype, n, d):
    x = wl.uniform(n, d, 7 + n + d)
    eps = wl.eps_quantile(x, [0.0, 0.01, 0.1, 0.5], 200, 3) if n > 1 else [0.0, 1.0]
    _check(ctx, x, eps, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_workload1_full(ctx, dtype):
    x, eps = wl.config(1)
    res = _check(ctx, x, eps, dtype)
    assert int(res.totals[0]) == int(res.counts[0].sum()) + x.shape[0]


@pytest.mark.parametrize("dtype", DTYPES)
def test_duplicates_and_eps_edges(ctx, dtype):
    rng = np.random.default_rng(11)
    base = rng.random((200, 24))
    x = np.concatenate([base, base[:50], base[:10]])  # rows 0..9 appear 3x, 10..49 twice
    eps = [0.0, -1.0, float("inf"), 1e300, 1e-300, 0.25]
    res = _check(ctx, x, eps, dtype)
    n = x.shape[0]
    # eps = 0: exactly the exact duplicates; eps < 0: nothing; inf / huge: everything.
    assert int(res.totals[1]) == 0 and res.counts[1].sum() == 0
    assert int(res.totals[2]) == n * n and int(res.totals[3]) == n * n
    assert int(res.totals[0]) == n + 2 * (10 * 3 + 40)


@pytest.mark.parametrize("dtype", DTYPES)
def test_exact_boundary_pairs(ctx, dtype):
    """eps set to the exact FP64 distance of real pairs: those pairs must count."""
    x = wl.gmm(3000, 64, 8, 0.05, 21)
    rng = np.random.default_rng(5)
    i = rng.integers(0, 3000, 6)
    j = rng.integers(0, 3000, 6)
    d2 = oracle.pair_d2(x, i, j)
    eps = [float(np.sqrt(v)) for v in d2]
    # sqrt is correctly rounded: sqrt(d2) <= eps holds with equality for these pairs.
    _check(ctx, x, eps, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_one_ulp_either_side_of_a_pair(ctx, dtype):
    """eps at a pair's exact FP64 distance and one ulp below it: the pair counts for the
    first and not for the second.  Every such pair falls inside the low-precision joins'
    band and inside the middle stage's radius, so only the FP64 refine can decide it."""
    x = wl.gmm(2500, 96, 8, 0.05, 23)
    rng = np.random.default_rng(8)
    i = rng.integers(0, 2500, 4)
    j = (i + 1 + rng.integers(0, 2499, 4)) % 2500
    eps = []
    for v in oracle.pair_d2(x, i, j):
        e = float(np.sqrt(v))
        eps += [e, float(np.nextafter(e, 0.0))]
    res = _check(ctx, x, eps, dtype)
    # The two eps really straddle the pair: its row holds it at eps (sqrt(d2) == eps) and
    # not one ulp below, so its count drops (counts are in the caller's eps order).
    for k in range(4):
        assert res.counts[2 * k, i[k]] > res.counts[2 * k + 1, i[k]]
    assert res.stats["fp64_pairs"] >= 1


@pytest.mark.parametrize("dtype", DTYPES)
def test_row_range_and_stride(ctx, dtype):
    x0 = wl.uniform(1700, 40, 9)
    wide = np.zeros((1700, 57))
    wide[:, :40] = x0
    x = wide[:, :40]  # ld = 57 > d
    eps = wl.eps_quantile(x0, [0.002, 0.05], 300, 1)
    _check(ctx, x, eps, dtype, row_begin=333, row_end=1201)


@pytest.mark.parametrize("dtype", ["fp64", "fp16"])
def test_workload2_reduced(ctx, dtype):
    x, eps = wl.config(2, n=20000)
    res = _check(ctx, x, eps, dtype)
    assert res.stats["band_pairs"] >= 0
    # The middle stage decides part of a low-precision band; FP64 passes refine all of it.
    assert 0 <= res.stats["fp64_pairs"] <= res.stats["band_pairs"]
    if dtype == "fp64":
        assert res.stats["fp64_pairs"] == res.stats["band_pairs"]
    elif res.stats["band_pairs"] > 100:
        assert res.stats["fp64_pairs"] < res.stats["band_pairs"]


@pytest.mark.parametrize("dtype", DTYPES)
def test_workload3_reduced(ctx, dtype):
    x, eps = wl.config(3, n=30000)
    _check(ctx, x, eps, dtype)


def _check_errors(ctx, res, dtype, tiles=2048):
    """bfj_pass_errors right after a pass: every candidate record and the sampled tiles
    lie inside the interval the band tests assumed, and the two bound ratios measured
    on the cast rows are at most 1 (the accumulation bound E_i and the quantisation
    radius q_i + q_j each hold on their own)."""
    e = ctx.pass_errors(max_candidates=0, sample_tiles=tiles, seed=7)
    assert e["violations"] == 0, e
    assert e["cand_total"] == res.stats["candidates"] and e["cand_pairs"] == e["cand_total"], e
    assert e["tile_pairs"] > 0
    if dtype == "fp64":
        assert e["max_acc_ratio"] == -1.0 and e["max_quant_ratio"] == -1.0
    else:
        assert 0.0 < e["max_acc_ratio"] <= 1.0, e
        assert 0.0 < e["max_quant_ratio"] <= 1.0, e
    return e


@pytest.mark.parametrize("workload,n", [(3, 20000), (4, 8000), (2, 10000)])
def test_pass_errors_within_the_certified_bounds(ctx, workload, n):
    """The max squared-distance error of every dtype, measured against the FP64 direct
    form, on the workloads' own data: FP64 is orders of magnitude closer than FP16/BF16,
    which are closer than FP8, and every dtype stays inside its certified interval."""
    x, eps = wl.config(workload, n=n)
    errs = {}
    for dtype in DTYPES:
        res = _check(ctx, x, eps, dtype)
        errs[dtype] = _check_errors(ctx, res, dtype)
    tile = {k: v["tile_max_abs_err"] for k, v in errs.items()}
    assert tile["fp64"] < 1e-6 * tile["fp16"], tile
    assert tile["fp16"] < tile["fp8"] and tile["bf16"] < tile["fp8"], tile
    assert tile["fp16"] > 0.0


def test_pass_errors_sampling_and_errors(ctx):
    x, eps = wl.config(3, n=6000)
    res = ctx.self_join(x, eps, dtype="fp16")
    full = ctx.pass_errors(max_candidates=0, sample_tiles=0)
    assert full["tile_pairs"] == 0 and full["cand_pairs"] == res.stats["candidates"]
    part = ctx.pass_errors(max_candidates=max(1, res.stats["candidates"] // 10), sample_tiles=16)
    assert 0 < part["cand_pairs"] < full["cand_pairs"] and part["cand_max_abs_err"] <= full["cand_max_abs_err"]
    # Another call on the context invalidates the measurement.
    ctx.reference_mode_join(x[:50], 0.1, 1, 64, 64)
    with pytest.raises(bfj.JoinError) as ei:
        ctx.pass_errors()
    assert ei.value.code == 1


@pytest.mark.parametrize("dtype", DTYPES)
def test_workload4_reduced(ctx, dtype):
    x, eps = wl.config(4, n=12000)
    _check(ctx, x, eps, dtype)


def test_multi_eps_equals_single(ctx):
    x, _ = wl.config(3, n=8000)
    eps = wl.eps_quantile(x, [0.0005, 0.001, 0.002, 0.005, 0.01, 0.02, 0.05, 0.1], 400, 2)
    multi = ctx.self_join(x, eps, dtype="fp16")
    for k, e in enumerate(eps):
        single = ctx.self_join(x, e, dtype="fp16")
        np.testing.assert_array_equal(multi.counts[k], single.counts[0])
        assert multi.totals[k] == single.totals[0]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("sym", [True, False])
def test_unsorted_eps_lists(ctx, dtype, sym):
    # Eps in no particular order, with repeats, a negative one and one beyond the
    # diameter: counts come back in the caller's order.
    x = wl.gmm(3000, 40, 6, 0.05, 91)
    q = wl.eps_quantile(x, [0.02, 0.001, 0.2], 300, 4)
    eps = [q[0], -1.0, q[1], 1e3, q[2], q[1], 0.0, q[0]]
    if sym:
        _check(ctx, x, eps, dtype)
    else:
        _check(ctx, x, eps, dtype, row_begin=700, row_end=2100)


@pytest.mark.parametrize("dtype", DTYPES)
def test_batch_mixed_dims(ctx, dtype):
    sets = wl.batch(n_sets=12, seed=77, n_min=100, n_max=3000)
    out, stats = ctx.self_join_batch(sets, dtype=dtype)
    for s, (counts, totals) in zip(sets, out):
        oc, ot = oracle.join(s["data"], s["eps"])
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(totals, ot)
    assert stats["pairs"] == sum(s["data"].shape[0] ** 2 for s in sets)


def test_plan_graph_replays(ctx):
    sets = wl.batch(n_sets=6, seed=78, n_min=200, n_max=2500)
    dev = [{"data": ctx.to_device(s["data"]), "eps": s["eps"],
            "counts": ctx.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
    want = [oracle.join(s["data"], s["eps"]) for s in sets]
    plan = ctx.plan(dev, dtype="fp16")
    # The shapes the plan tuned for its K-chunk groups (on its own datasets) are built ones.
    tiles = plan.tiles()
    assert tiles and len({t["kc"] for t in tiles}) == len(tiles)
    for t in tiles:
        assert t["rows"] in (16, 32, 64, 96, 128) and t["blocks_per_cu"] >= 1
        assert t["panel_bytes"] in (1 << 19, 1 << 20, 1 << 21)
    for rep in range(3):
        plan.run()
        for dv, tot, (oc, ot) in zip(dev, plan.totals(), want):
            np.testing.assert_array_equal(dv["counts"].numpy(), oc, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(tot, ot)
        # eager work on the same context between replays
        ctx.self_join(sets[0]["data"], sets[0]["eps"], dtype="fp64")
    plan.close()


def test_items_of_a_destroyed_plan_are_never_reused(ctx):
    """A plan's work items live as long as the plan.  After it is destroyed, an eager
    call with a new shape must upload its own items, whatever address its item list
    gets (it used to be recognised by address, and a new list at the freed address was
    taken for the plan's)."""
    sets = wl.batch(n_sets=3, seed=79, n_min=300, n_max=1500)
    dev = [{"data": ctx.to_device(s["data"]), "eps": s["eps"],
            "counts": ctx.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
    xa = wl.uniform(2100, 24, 1)
    xb = wl.gmm(3300, 40, 5, 0.05, 2)
    ea = wl.eps_quantile(xa, [0.01], 200, 1)
    eb = wl.eps_quantile(xb, [0.005, 0.02], 200, 2)
    for rep in range(3):
        plan = ctx.plan(dev, dtype="fp16")
        _check(ctx, xa, ea, "fp16")
        plan.run()
        for dv, s in zip(dev, sets):
            np.testing.assert_array_equal(dv["counts"].numpy(), oracle.join(s["data"], s["eps"])[0])
        plan.close()
        _check(ctx, xb[: 3300 - 100 * rep], eb, "fp16")


def test_column_slice_at_the_end_of_its_buffer(tmp_path):
    """A column slice of a wider array whose last row ends where the buffer ends, with
    an unreadable page right after it: the host-buffer entry points copy exactly
    (n - 1) ld + d values (they used to copy n ld from the slice's start, past the end).
    Run in a child process, so that a read past the end fails that process only."""
    import subprocess
    import sys
    code = f"""
import ctypes, mmap, sys
import numpy as np
sys.path.insert(0, {ROOT_DIR!r})
import tensor_bf_main_amd as bfj
from oracle import oracle
n, ld, d0, d = 900, 50, 10, 40
page = mmap.PAGESIZE
nbytes = n * ld * 8
total = ((nbytes + page - 1) // page + 1) * page
m = mmap.mmap(-1, total)
base = ctypes.addressof(ctypes.c_char.from_buffer(m))
libc = ctypes.CDLL(None)
libc.mprotect.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
assert libc.mprotect(base + total - page, page, 0) == 0  # PROT_NONE after the array
wide = np.frombuffer(m, dtype=np.float64, count=n * ld, offset=total - page - nbytes).reshape(n, ld)
wide[:] = np.random.default_rng(3).random((n, ld))
x = wide[:, d0:]
eps = [0.9, 1.2]
with bfj.Context(0) as ctx:
    for dt in ("fp16", "fp64"):
        res = ctx.self_join(x, eps, dtype=dt)
        oc, ot = oracle.join(x, eps)
        assert np.array_equal(res.counts, oc) and np.array_equal(res.totals, ot), dt
        pairs, tot, _ = ctx.self_join_pairs(np.ascontiguousarray(x), eps[0], dtype=dt)
        assert tot == int(ot[0])
print("ok")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.returncode, r.stderr[-2000:])


@pytest.mark.parametrize("dtype", ["fp16", "fp64", "fp8"])
def test_forced_overflow_grows_and_repeats(ctx, dtype):
    """Capacities far below what a pass needs: the candidate stream and the band list
    overflow, the library grows both to what the pass reported and repeats it; counts
    are exact and band_retries says it happened.  Eager calls and a captured plan."""
    x, eps = wl.config(3, n=9000)
    ctx.set_capacity(candidate_chunks=2, band_records=16)
    res = _check(ctx, x, eps, dtype)
    assert res.stats["band_retries"] >= 1
    assert res.stats["candidates"] > 2 * 1024
    # The grown capacities stay: the next pass runs once.
    res = _check(ctx, x, eps, dtype)
    assert res.stats["band_retries"] == 0
    sets = wl.batch(n_sets=4, seed=81, n_min=1000, n_max=3000)
    dev = [{"data": ctx.to_device(s["data"]), "eps": s["eps"],
            "counts": ctx.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
    plan = ctx.plan(dev, dtype=dtype)
    ctx.set_capacity(candidate_chunks=1, band_records=8)
    st = plan.run()
    assert st["band_retries"] >= 1
    for dv, s, tot in zip(dev, sets, plan.totals()):
        oc, ot = oracle.join(s["data"], s["eps"])
        np.testing.assert_array_equal(dv["counts"].numpy(), oc)
        np.testing.assert_array_equal(tot, ot)
    assert plan.run()["band_retries"] == 0
    plan.close()


def _adversarial(kind, seed):
    """Data built to stress the error bound (each is checked for every dtype against the
    oracle, with eps at the FP64 distances of real pairs)."""
    rng = np.random.default_rng(seed)
    if kind == "offset":            # a large common offset: centring must not lose the distances
        x = 1e6 + rng.random((2500, 32))
    elif kind == "row_scales":      # row norms spread over 2^-40 .. 2^40
        x = rng.standard_normal((2500, 48)) * np.exp2(rng.uniform(-40, 40, (2500, 1)))
    elif kind == "subnormal":       # all but a few rows subnormal in FP16 / e4m3 after the scale
        x = rng.random((2500, 40)) * 1e-7
        x[:5] = rng.random((5, 40))
    elif kind == "dups_1e-12":      # duplicates with 1e-12 jitter: distances far below any cast's resolution
        base = rng.random((1200, 64))
        x = np.concatenate([base, base + rng.standard_normal(base.shape) * 1e-12,
                            base[:300] + rng.standard_normal((300, 64)) * 1e-12])
    elif kind == "cancel_d256":     # rows that nearly cancel: tiny distances between large rows
        c = rng.random((1, 256)) * 100.0
        x = c + rng.standard_normal((2000, 256)) * 1e-4
    else:
        raise ValueError(kind)
    n = x.shape[0]
    # eps at the exact FP64 distance of each of three rows to its nearest other row (the
    # pairs then lie inside every cast's band), and two quantiles
    exact = []
    for i in rng.integers(0, n, 3):
        d2 = oracle.pair_d2(x, np.full(n, i), np.arange(n))
        d2[i] = np.inf
        exact.append(float(np.sqrt(d2.min())))
    return x, exact + wl.eps_quantile(x, [0.0005, 0.01], 300, seed)


@pytest.mark.parametrize("kind", ["offset", "row_scales", "subnormal", "dups_1e-12", "cancel_d256"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_adversarial_data_stays_exact(ctx, kind, dtype):
    x, eps = _adversarial(kind, 100 + len(kind))
    res = _check(ctx, x, eps, dtype)
    e = ctx.pass_errors(max_candidates=0, sample_tiles=512)
    assert e["violations"] == 0, e
    if dtype != "fp64" and e["tile_pairs"]:
        assert e["max_acc_ratio"] <= 1.0 and e["max_quant_ratio"] <= 1.0, e
    # the first three eps are distances of real pairs: each holds at least that pair
    assert all(int(res.totals[k]) >= x.shape[0] + 2 for k in range(3))


def test_allocation_failure_is_reported_and_the_context_recovers():
    """Every device allocation of a context's first pass, failed in turn (the k-th one
    from the start of the call): the call returns BFJ_ERR_NOMEM with the reason, and the
    same context then runs the pass correctly.  The same for a plan's creation."""
    import tensor_bf_main_amd as bfj_
    lib = bfj_.load_library()
    x, eps = wl.config(3, n=4000)
    oc, ot = oracle.join(x, eps)
    failures = 0
    try:
        for k in range(1, 80):
            with bfj_.Context(0) as c:
                assert lib.bfj_fail_allocation(k) == 0
                try:
                    c.self_join(x, eps, dtype="fp16")
                    hit = False
                except bfj_.JoinError as err:
                    assert err.code == 3 and "injected" in str(err), str(err)
                    hit = True
                lib.bfj_fail_allocation(0)
                res = c.self_join(x, eps, dtype="fp16")
                np.testing.assert_array_equal(res.counts, oc)
                np.testing.assert_array_equal(res.totals, ot)
                failures += hit
            if not hit:
                break
        assert failures >= 20, failures  # a first pass allocates every buffer of the context
        sets = wl.batch(n_sets=3, seed=82, n_min=500, n_max=1500)
        with bfj_.Context(0) as c:
            dev = [{"data": c.to_device(s["data"]), "eps": s["eps"],
                    "counts": c.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
            lib.bfj_fail_allocation(3)
            with pytest.raises(bfj_.JoinError) as ei:
                c.plan(dev, dtype="fp16")
            assert ei.value.code == 3
            lib.bfj_fail_allocation(0)
            plan = c.plan(dev, dtype="fp16")
            plan.run()
            for dv, s in zip(dev, sets):
                np.testing.assert_array_equal(dv["counts"].numpy(), oracle.join(s["data"], s["eps"])[0])
            plan.close()
    finally:
        lib.bfj_fail_allocation(0)


def test_device_resident_path(ctx):
    x, eps = wl.config(1, n=5000)
    xd = ctx.to_device(x)
    cnt = ctx.empty((1, 5000), np.int64)
    res = ctx.self_join_device(xd, eps, dtype="fp16", counts=cnt)
    oc, ot = oracle.join(x, eps)
    np.testing.assert_array_equal(cnt.numpy(), oc)
    np.testing.assert_array_equal(res.totals, ot)
    # row shard of the same device array
    cs = ctx.empty((1, 1000), np.int64)
    ctx.self_join_device(xd, eps, dtype="fp64", row_begin=2000, row_end=3000, counts=cs)
    np.testing.assert_array_equal(cs.numpy(), oc[:, 2000:3000])


def test_nonfinite_rows_rejected(ctx):
    x = wl.uniform(100, 8, 1)
    x[17, 3] = np.nan
    with pytest.raises(bfj.JoinError) as ei:
        ctx.self_join(x, 0.5, dtype="fp16")
    assert ei.value.code == 1 and "NaN" in str(ei.value)
    x[17, 3] = np.inf
    with pytest.raises(bfj.JoinError):
        ctx.self_join(x, 0.5, dtype="fp64")
    # the context stays usable
    _check(ctx, wl.uniform(100, 8, 2), 0.5, "fp16")


def test_autotune_then_exact(ctx):
    r = ctx.autotune("fp16", 128, n=8192)
    assert r in bfj.tile_rows_options("fp16", 128)
    x, eps = wl.config(3, n=6000)
    _check(ctx, x, eps, "fp16")
    ctx.set_tile_rows("fp16", 0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("d", [16, 96, 200, 256])
def test_every_built_tile_shape_is_exact(ctx, dtype, d):
    """Every strip width built for (dtype, d), forced in turn, on the symmetric sweep and on
    a row range: the oracle's counts.  Clustered rows and eps up to the 2 % quantile make hot
    tiles common, and n = 1507 leaves a ragged last block and diagonal blocks in every strip."""
    x = wl.gmm(1507, d, 8, 0.05, 31 + d)
    eps = wl.eps_quantile(x, [0.002, 0.02], 300, 5)
    oc, ot = oracle.join(x, eps)
    oc_r, ot_r = oracle.join(x, eps, 100, 1200)
    try:
        for r in bfj.tile_rows_options(dtype, d):
            ctx.set_tile_rows(dtype, r)
            res = ctx.self_join(x, eps, dtype=dtype)
            np.testing.assert_array_equal(res.counts, oc, err_msg=f"{dtype} d={d} R={r}")
            np.testing.assert_array_equal(res.totals, ot, err_msg=f"{dtype} d={d} R={r}")
            res = ctx.self_join(x, eps, dtype=dtype, row_begin=100, row_end=1200)
            np.testing.assert_array_equal(res.counts, oc_r, err_msg=f"{dtype} d={d} R={r} rows 100-1200")
            np.testing.assert_array_equal(res.totals, ot_r, err_msg=f"{dtype} d={d} R={r} rows 100-1200")
    finally:
        ctx.set_tile_rows(dtype, 0)


@pytest.mark.slow
def test_workload3_full_fp16_vs_fp64(ctx):
    """n = 1M, d = 128: FP16+refine and FP64 counts agree exactly; 1000 seeded rows
    match the oracle; the ordered-pair relation is symmetric (even total)."""
    x, eps = wl.config(3)
    n = x.shape[0]
    xd = ctx.to_device(x)
    c16 = ctx.empty((1, n), np.int64)
    c64 = ctx.empty((1, n), np.int64)
    r16 = ctx.self_join_device(xd, eps, dtype="fp16", counts=c16)
    r64 = ctx.self_join_device(xd, eps, dtype="fp64", counts=c64)
    a16, a64 = c16.numpy()[0], c64.numpy()[0]
    np.testing.assert_array_equal(a16, a64)
    assert r16.totals[0] == r64.totals[0]
    assert int(a16.sum()) % 2 == 0
    rows = np.random.default_rng(4242).choice(n, 1000, replace=False)
    oc = oracle.join_rows(x, rows, eps)[0]
    np.testing.assert_array_equal(a16[rows], oc)


@pytest.mark.slow
def test_workload2_full_fp64_and_fp16():
    """n = 100k, d = 128 uniform: the FP64-MFMA and FP16 paths agree on every per-point
    count, 1000 seeded rows match the oracle, and both stay inside their certified error
    intervals."""
    import tensor_bf_main_amd as bfj_
    x, eps = wl.config(2)
    n = x.shape[0]
    with bfj_.Context(0) as c:
        xd = c.to_device(x)
        got = {}
        for dtype in ("fp64", "fp16"):
            cnt = c.empty((1, n), np.int64)
            res = c.self_join_device(xd, eps, dtype=dtype, counts=cnt)
            e = c.pass_errors(max_candidates=1 << 22, sample_tiles=1024)
            assert e["violations"] == 0, (dtype, e)
            got[dtype] = (cnt.numpy()[0], res.totals[0])
        np.testing.assert_array_equal(got["fp16"][0], got["fp64"][0])
        assert got["fp16"][1] == got["fp64"][1]
        rows = np.random.default_rng(4343).choice(n, 1000, replace=False)
        np.testing.assert_array_equal(got["fp64"][0][rows], oracle.join_rows(x, rows, eps)[0])
        # about 64 neighbours per point, as the workload asks
        assert 32 < (int(got["fp64"][1]) - n) / n < 128


@pytest.mark.slow
def test_workload4_full_bf16_fp8_vs_fp64():
    """n = 500k, d = 256 unit-norm rows with 5 % near-duplicates: BF16 and FP8 (per-row
    scales) each give the FP64-MFMA join's per-point counts exactly; 500 seeded rows
    match the oracle; no pair leaves its certified interval."""
    import tensor_bf_main_amd as bfj_
    x, eps = wl.config(4)
    n = x.shape[0]
    with bfj_.Context(0) as c:
        xd = c.to_device(x)
        cnt = c.empty((1, n), np.int64)
        r64 = c.self_join_device(xd, eps, dtype="fp64", counts=cnt)
        want = cnt.numpy()[0]
        rows = np.random.default_rng(4444).choice(n, 500, replace=False)
        np.testing.assert_array_equal(want[rows], oracle.join_rows(x, rows, eps)[0])
        for dtype in ("bf16", "fp8"):
            res = c.self_join_device(xd, eps, dtype=dtype, counts=cnt)
            e = c.pass_errors(max_candidates=1 << 22, sample_tiles=1024)
            assert e["violations"] == 0, (dtype, e)
            np.testing.assert_array_equal(cnt.numpy()[0], want, err_msg=dtype)
            assert res.totals[0] == r64.totals[0]
            assert res.stats["fp64_pairs"] <= res.stats["band_pairs"]


@pytest.mark.slow
def test_workload5_full_graph_vs_eager():
    """512 datasets x 4 eps in one captured plan: every replay gives the per-dataset
    eager counts, and a seeded few whole datasets match the oracle."""
    import tensor_bf_main_amd as bfj_
    sets = wl.batch(n_sets=512)
    with bfj_.Context(0) as c:
        dev = [{"data": c.to_device(s["data"]), "eps": s["eps"],
                "counts": c.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
        plan = c.plan(dev, dtype="fp16")
        plan.run()
        e = c.pass_errors(max_candidates=1 << 22, sample_tiles=2048)
        assert e["violations"] == 0, e
        st = plan.run()
        assert st["pairs"] == sum(s["data"].shape[0] ** 2 for s in sets)
        graph = [dv["counts"].numpy() for dv in dev]
        gtot = plan.totals()
        plan.close()
        for k, (dv, s) in enumerate(zip(dev, sets)):
            res = c.self_join_device(dv["data"], s["eps"], dtype="fp16", counts=dv["counts"])
            np.testing.assert_array_equal(dv["counts"].numpy(), graph[k], err_msg=f"set {k}")
            np.testing.assert_array_equal(res.totals, gtot[k])
        small = [k for k, s in enumerate(sets) if s["data"].shape[0] <= 6000]
        for k in np.random.default_rng(5).choice(small, 6, replace=False):
            oc, ot = oracle.join(sets[k]["data"], sets[k]["eps"])
            np.testing.assert_array_equal(graph[k], oc, err_msg=f"set {k}")
            np.testing.assert_array_equal(gtot[k], ot)


@pytest.mark.parametrize("mode,dtype", [(1, "fp64"), (2, "fp64"), (3, "fp16"), (4, "fp16"), (7, "fp64"),
                                        (8, "fp16"), (9, "bf16"), (10, "fp8")])
def test_cli_drop_in(tmp_path, mode, dtype):
    """bfjoin_main keeps the reference CLI (main.cu:16-128): same arguments, console
    lines and CSV log; every searchMode returns the exact total."""
    import subprocess

    gold = json.load(open(os.path.join(GOLDEN, "expected_counts.json")))["test128.txt"]
    k = gold["eps"].index(0.05)
    log = tmp_path / "log.csv"
    counts = tmp_path / "counts.txt"
    pairs = tmp_path / "pairs.txt"
    r = subprocess.run([bfj.CLI_PATH, os.path.join(GOLDEN, "test128.txt"), "0.05", str(mode), "0", "--dim", "18",
                        "--log", str(log), "--counts", str(counts), "--pairs", str(pairs)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"[Main | Result] ~ Total result set size: {gold['totals'][k]}" in r.stdout
    # The result set: (total - n) / 2 unordered pairs, i < j, no repeats.
    got = [tuple(int(v) for v in ln.split()) for ln in pairs.read_text().splitlines()]
    assert len(got) == len(set(got)) == (gold["totals"][k] - 128) // 2 and all(i < j for i, j in got)
    rows = log.read_text().splitlines()
    assert rows[0].startswith("Dataset, epsilon, searchMode, executionTime, totalNeighbors, inputDim, computeDim")
    f = [v.strip() for v in rows[1].split(",")]
    assert int(f[2]) == mode and int(f[4]) == gold["totals"][k] and f[11] == dtype
    assert [int(v) for v in counts.read_text().split()] == gold["counts"][k]


def _oracle_pairs(x, eps):
    """Every unordered pair i < j with d2 <= T(eps), from the oracle's direct form."""
    n = x.shape[0]
    iu, ju = np.triu_indices(n, 1)
    d2 = oracle.pair_d2(x, iu, ju)
    keep = d2 <= oracle.threshold(eps)
    return set(zip(iu[keep].tolist(), ju[keep].tolist()))


@pytest.mark.parametrize("dtype", DTYPES)
def test_pair_list_matches_oracle(ctx, dtype):
    x = wl.gmm(1800, 24, 8, 0.05, 31)
    for q in (0.002, 0.03):
        eps = float(wl.eps_quantile(x, [q], 300, 5)[0])
        pairs, total, st = ctx.self_join_pairs(x, eps, dtype=dtype)
        got = set(map(tuple, pairs.astype(np.int64).tolist()))
        assert len(got) == pairs.shape[0], "duplicate pairs"
        assert all(i < j for i, j in got)
        assert got == _oracle_pairs(x, eps)
        assert total == x.shape[0] + 2 * pairs.shape[0]
        res = ctx.self_join(x, [eps], dtype=dtype)
        assert int(res.totals[0]) == total


@pytest.mark.parametrize("dtype", ["fp64", "fp16"])
def test_pair_list_edges(ctx, dtype):
    x = wl.uniform(700, 9, 77)
    x[350:360] = x[0]  # exact duplicates: distance 0 pairs
    # eps < 0: no pair; eps = 0: only the duplicates; eps beyond the diameter: all pairs.
    pairs, total, _ = ctx.self_join_pairs(x, -1.0, dtype=dtype)
    assert pairs.shape[0] == 0 and total == 0
    pairs, total, _ = ctx.self_join_pairs(x, 0.0, dtype=dtype)
    assert set(map(tuple, pairs.astype(np.int64).tolist())) == _oracle_pairs(x, 0.0)
    pairs, total, _ = ctx.self_join_pairs(x, 1e6, dtype=dtype, capacity=10)  # grown after CAPACITY
    n = x.shape[0]
    assert pairs.shape[0] == n * (n - 1) // 2 and total == n * n
    iu, ju = np.triu_indices(n, 1)
    np.testing.assert_array_equal(pairs[:, 0], iu)
    np.testing.assert_array_equal(pairs[:, 1], ju)
    # The raw call with a small buffer reports CAPACITY and the number of pairs it needs.
    import ctypes
    xc = np.ascontiguousarray(x)
    buf = np.zeros((5, 2), dtype=np.uint32)
    found, tot = ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = ctx.lib.bfj_self_join_pairs(ctx._h, xc.ctypes.data, n, xc.shape[1], xc.shape[1], 1e6, bfj.DTYPES[dtype],
                                     buf.ctypes.data, 5, ctypes.byref(found), ctypes.byref(tot), None)
    assert rc == bfj.BFJ_ERR_CAPACITY and found.value == n * (n - 1) // 2


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,d,kind", [(512, 18, "gmm"), (3000, 128, "gmm"), (2500, 33, "uniform"), (4097, 200, "gmm")])
def test_locality_order_gives_the_same_results(ctx, dtype, n, d, kind):
    """Whole-dataset passes pack rows grouped by nearest pivot (order.hip); per-point
    counts, totals and pair lists must not depend on it, and must match the oracle."""
    x = wl.gmm(n, d, 8, 0.05, n + d) if kind == "gmm" else wl.uniform(n, d, n + d)
    eps = wl.eps_quantile(x, [0.002, 0.02], 200, 5)
    try:
        ctx.set_reorder(False)
        plain = ctx.self_join(x, eps, dtype=dtype)
        plain_pairs, _, _ = ctx.self_join_pairs(x, eps[0], dtype=dtype)
    finally:
        ctx.set_reorder(True)
    res = _check(ctx, x, eps, dtype)
    np.testing.assert_array_equal(res.counts, plain.counts)
    np.testing.assert_array_equal(res.totals, plain.totals)
    pairs, total, _ = ctx.self_join_pairs(x, eps[0], dtype=dtype)
    assert total == int(res.totals[0])
    assert np.all(pairs[:, 0] < pairs[:, 1])
    key = lambda p: np.sort(p[:, 0].astype(np.int64) * n + p[:, 1])
    np.testing.assert_array_equal(key(pairs), key(plain_pairs))


def test_argument_errors_on_a_live_context(ctx):
    """Bad arguments fail with BFJ_ERR_INVALID / BFJ_ERR_UNSUPPORTED and a message that
    names them, before any device work, and leave the context usable; n = 0 is a valid
    (empty) join."""
    lib = ctx.lib
    x = np.random.default_rng(5).random((64, 300))
    eps1 = (ctypes.c_double * 1)(0.5)
    tot = (ctypes.c_uint64 * 9)()

    def call(n, d, ld, eps, n_eps, dtype=1, rb=0, re=None, data=True):
        return lib.bfj_self_join(ctx._h, x.ctypes.data if data else None, n, d, ld, eps, n_eps, dtype, rb,
                                 n if re is None else re, None, tot, None)

    cases = [
        (dict(n=64, d=0, ld=300, eps=eps1, n_eps=1), 1, b"d <= ld"),
        (dict(n=64, d=257, ld=300, eps=eps1, n_eps=1), 4, b"257"),
        (dict(n=64, d=20, ld=10, eps=eps1, n_eps=1), 1, b"d <= ld"),
        (dict(n=-1, d=20, ld=300, eps=eps1, n_eps=1), 1, b"n < 0"),
        (dict(n=64, d=20, ld=300, eps=(ctypes.c_double * 9)(*[0.1] * 9), n_eps=9), 1, b"n_eps"),
        (dict(n=64, d=20, ld=300, eps=(ctypes.c_double * 1)(float("nan")), n_eps=1), 1, b"NaN"),
        (dict(n=64, d=20, ld=300, eps=eps1, n_eps=1, rb=10, re=5), 1, b"row range"),
        (dict(n=64, d=20, ld=300, eps=eps1, n_eps=1, re=65), 1, b"row range"),
        (dict(n=64, d=20, ld=300, eps=eps1, n_eps=1, dtype=9), 1, b"dtype"),
        (dict(n=64, d=20, ld=300, eps=eps1, n_eps=1, data=False), 1, b"NULL"),
    ]
    for kw, code, word in cases:
        assert call(**kw) == code, kw
        assert word in lib.bfj_last_error(), (kw, lib.bfj_last_error())
    # An empty dataset joins to nothing.
    assert call(n=0, d=20, ld=300, eps=eps1, n_eps=1, data=False) == 0
    assert tot[0] == 0
    # The context still works after the errors.
    _check(ctx, x[:, :20], 0.5, "fp16")
