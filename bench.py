"""Benchmark: brute-force Euclidean epsilon self-join on one MI355X.

Headline (default, --workload 3): n = 1,000,000 points, d = 128, Gaussian mixture
(64 centres uniform in [0,1)^128, sigma 0.03 per dimension), eps set for a mean of
~100 neighbours.  One step = one full pass of the join (centre/cast/norms/
thresholds, MFMA join with fused epsilon epilogue, FP64 band refine, count
finalisation) over the dataset already resident in HBM.  The FP16-MFMA path is
the measured value; the plain FP64-MFMA join runs on the same data for the
speedup, and the two must give identical per-point counts (plus seeded rows
against the CPU oracle).

The headline run also measures the other configurations at their stated sizes and
adds each as a sub-record under "configs" (--no-sub-configs: headline only).  Every
record carries "errors": max |d2_approx - d2_fp64| over the pass's candidate pairs and
over seeded tiles of all pairs, checked against the certified error interval.

Other workloads alone (one JSON line each, same keys):
  --workload 2   n=100,000 d=128 uniform, FP64-MFMA join (exact vs oracle rows)
  --workload 4   n=500,000 d=256 unit-norm rows with 5% near-duplicates, --dtype bf16|fp8|fp16,
                 compared with the FP64-MFMA join
  --workload 5   512 datasets (n log-uniform in [1000, 65536], d in {16,64,128,200}),
                 4 eps each, FP16: one persistent hipGraph launch vs per-dataset launches

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 3] [--n N] ...
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp16": 2500.0, "bf16": 2500.0, "fp8": 5000.0, "fp64": 78.6}  # MI355X dense MFMA


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload, dtype, tile_rows):
    """HBM bytes per join launch from the committed rocprofv3 PMC summaries
    (FETCH_SIZE and WRITE_SIZE runs scaled by a copy of known size), for the
    configuration they were measured on; None otherwise."""
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    want = f"workload={workload} dtype={dtype} tile_rows={tile_rows} "
    if not os.path.isdir(prof):
        return None
    for name in sorted(os.listdir(prof)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(prof, name)):
            if line.startswith("machine-readable: " + want):
                return float(line.split("traffic_bytes=")[1])
    return None


def roofline(flops, join_ms, dtype, traffic=None):
    achieved = flops / (join_ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[dtype], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS[dtype], "traffic": traffic,
            "kernel": "k_join (MFMA join + fused epsilon epilogue)"}


def host_cpu():
    """CPU model, logical cores of the host and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cores": os.cpu_count(), "usable_cores": usable}


def cpu_baseline(x, eps, rows_n, got_counts, seed=4242):
    """Oracle on a bounded sample of query rows; returns (baseline dict, mismatches)."""
    from oracle import oracle

    n = x.shape[0]
    rows = np.random.default_rng(seed).choice(n, rows_n, replace=False)
    log(f"cpu oracle on {rows_n} rows x {n} candidates, {oracle.num_threads()} threads")
    t0 = time.perf_counter()
    oc = oracle.join_rows(x, rows, eps)
    tc = time.perf_counter() - t0
    mism = int((got_counts[:, rows] != oc).sum())
    what = f"all {n} rows (full run)" if rows_n == n else f"{rows_n} seeded query rows"
    return {"value": rows_n * float(n) / tc, "unit": "pair-distances/s", "cores": oracle.num_threads(),
            "kind": "port", "sample": f"{what} x all {n} candidates, FP64 direct form "
                                      f"(OpenMP), {tc:.3g} s", "host": host_cpu()}, mism


def pass_errors(ctx, max_candidates=1 << 24, sample_tiles=4096):
    """Accuracy of the pass just run (outside the timed region): max |d2_approx - d2_fp64|
    over its candidate records (every k-th chunk beyond max_candidates) and over seeded
    tiles of all pairs, the two bound ratios, and the pairs outside the certified interval."""
    e = ctx.pass_errors(max_candidates=max_candidates, sample_tiles=sample_tiles, seed=2024)
    return {"max_abs_d2_err_candidates": e["cand_max_abs_err"], "max_abs_d2_err_sampled": e["tile_max_abs_err"],
            "candidates_checked": int(e["cand_pairs"]), "candidates_total": int(e["cand_total"]),
            "sampled_pairs": int(e["tile_pairs"]), "bound_violations": int(e["violations"]),
            "max_accumulation_bound_ratio": e["max_acc_ratio"], "max_quantisation_bound_ratio": e["max_quant_ratio"]}


def single_dataset(args, ctx, x, eps, dtype, metric, workload, fp64_counts=None):
    """One dataset: timed steps of ``dtype``, then (untimed) the accuracy statistics, the
    host-buffer rate, the FP64-MFMA join on the same rows (or ``fp64_counts`` from an
    earlier run of it) and the CPU oracle on a row sample."""
    n, d = x.shape
    e = np.atleast_1d(eps)
    xd = ctx.to_device(x)  # rows resident in HBM before the timed region
    c_main = ctx.empty((e.size, n), np.int64)

    def step(dt, counts):
        # Each call enqueues the whole pass on the library stream and returns after
        # that stream has synchronised, so the wall clock brackets complete steps.
        return ctx.self_join_device(xd, e, dtype=dt, counts=counts)

    log(f"warmup {args.warmup} x {dtype}")
    for _ in range(args.warmup):
        step(dtype, c_main)
    join_ms, kern_ms, last = [], [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        last = step(dtype, c_main)
        join_ms.append(last.stats["ms_join"])
        kern_ms.append(last.stats["ms_total"])
    t_main = (time.perf_counter() - t0) / args.steps
    pairs = float(n) * float(n)
    log(f"{dtype}: {t_main * 1e3:.1f} ms/step, join kernel {np.mean(join_ms):.1f} ms, "
        f"refine {last.stats['ms_refine']:.1f} ms, band {last.stats['band_pairs']}, FP64 {last.stats['fp64_pairs']}")
    out = {
        "metric": metric, "value": pairs / t_main, "unit": "pair-distances/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_main * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": workload, "n": n, "d": d, "eps": [float(v) for v in e],
                   "mean_neighbours": float(last.totals[0] - n) / n, "path": f"{dtype}-mfma+fp64-refine"},
        "total_pairs_within_eps": [int(t) for t in last.totals],
        "band_pairs": int(last.stats["band_pairs"]),
        "band_fraction": last.stats["band_pairs"] / pairs,
        "fp64_refine_pairs": int(last.stats["fp64_pairs"]),
        "candidates": int(last.stats["candidates"]),
        "slices": int(last.stats["slices"]),
        "max_band_halfwidth": last.stats["max_band_halfwidth"],
        "stage_ms": {k: last.stats[k] for k in ("ms_prep", "ms_join", "ms_refine", "ms_total")},
        # Kernel-only time of a pass (prep + join + classify/refine/finalize, HIP events),
        # median over the timed steps; ms_per_step adds the host side of each call.
        "kernel_ms_median": float(np.median(kern_ms)),
        "tile_rows": last.stats["tile_rows"],
    }
    # Algorithmic flops: true d, FMA = 2, one dot product per unordered pair (the
    # symmetric sweep decides both (i, j) and (j, i) from it).
    flops = 2.0 * d * n * (n + 1) / 2
    out["roofline"] = roofline(flops, float(np.mean(join_ms)), dtype,
                               None if args.n else measured_traffic(args.workload, dtype, int(last.stats["tile_rows"])))
    out["errors"] = pass_errors(ctx)
    got = c_main.numpy()
    # The host-buffer entry point (rows in host memory, one H2D copy per call, counts
    # copied back): the PCIe-inclusive rate, reported beside the value, never as it.
    t0 = time.perf_counter()
    for _ in range(2):
        host_res = ctx.self_join(x, e, dtype=dtype)
    t_host = (time.perf_counter() - t0) / 2
    out["host_buffers_ms_per_step"] = t_host * 1e3
    out["host_buffers_pairs_per_s"] = pairs / t_host
    host_mism = int((host_res.counts != got).sum())
    mismatches = host_mism
    out["fp64_counts"] = None
    if dtype != "fp64" and fp64_counts is not None:
        mismatches += int((got != fp64_counts).sum())
    elif dtype != "fp64" and not args.skip_fp64:
        c64 = ctx.empty((e.size, n), np.int64)
        log("fp64 warmup")
        step("fp64", c64)
        t0 = time.perf_counter()
        r64 = None
        for _ in range(args.fp64_steps):
            r64 = step("fp64", c64)
        t64 = (time.perf_counter() - t0) / args.fp64_steps
        out["fp64_counts"] = c64.numpy()
        mismatches += int((got != out["fp64_counts"]).sum()) + int((r64.totals != last.totals).sum())
        out["fp64_ms_per_step"] = t64 * 1e3
        out["fp64_pairs_per_s"] = pairs / t64
        out["speedup_vs_fp64"] = t64 / t_main
        out["fp64_join_tflops"] = flops / (r64.stats["ms_join"] * 1e-3) / 1e12
        log(f"fp64: {t64 * 1e3:.1f} ms/step; speedup {t64 / t_main:.2f}x; mismatches {mismatches}")
    if not args.skip_cpu:
        # About 10-30 s of CPU work; the whole dataset when that is less (n = 10k).
        rows_n = min(n, args.cpu_rows or max(64, min(12000, int(6.0e9 / n * 128 / d))))
        out["cpu_baseline"], sm = cpu_baseline(x, e, rows_n, got)
        out["oracle_sample_mismatches"] = sm
        mismatches = (mismatches or 0) + sm
    out["mismatches"] = mismatches
    xd.free()
    c_main.free()
    return out


def sub_configs(args, ctx):
    """The other configurations at their full sizes, each a sub-record of the headline
    line with the same measurements (timed steps, stage times, roofline, accuracy,
    mismatches against the FP64 path and the CPU oracle on a row sample)."""
    from tensor_bf_main_amd import workloads as wl

    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.n, sub.cpu_rows, sub.skip_cpu = args.sub_steps, 1, 0, 0, False
    out = {}

    def run(key, fn):
        t0 = time.perf_counter()
        try:
            rec = fn()
        except Exception as exc:  # a failing sub-record must not hide the headline
            rec = {"error": f"{type(exc).__name__}: {exc}"}
        rec.pop("fp64_counts", None)
        rec["wall_s"] = time.perf_counter() - t0
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data", "unit"):
            rec.pop(k, None)
        out[key] = rec
        log(f"{key}: {json.dumps({k: rec.get(k) for k in ('ms_per_step', 'value', 'mismatches', 'error')})}")

    def c1():
        x, eps = wl.config(1)
        sub.workload = 1
        return single_dataset(sub, ctx, x, eps, "fp64", "", "uniform_n10k_d32")

    def c2():
        x, eps = wl.config(2)
        sub.workload, sub.cpu_rows = 2, 2000
        return single_dataset(sub, ctx, x, eps, "fp64", "", "uniform")

    x4, e4 = wl.config(4)
    keep = {}

    def c4(dtype):
        def f():
            sub.workload, sub.cpu_rows, sub.skip_fp64 = 4, 400, False
            rec = single_dataset(sub, ctx, x4, e4, dtype, "", "unit_norm_5pct_dups", keep.get("fp64"))
            if rec.get("fp64_counts") is not None:
                keep["fp64"] = rec["fp64_counts"]
            return rec
        return f

    run("1_fp64_n10k_d32", c1)
    run("2_fp64_n100k_d128", c2)
    run("4_bf16_n500k_d256", c4("bf16"))
    run("4_fp8_n500k_d256", c4("fp8"))
    del x4, keep

    def c5():
        sub.workload, sub.dtype = 5, "fp16"
        return batch_workload(sub, ctx)

    run("5_fp16_batch512", c5)
    return out


def batch_workload(args, ctx):
    from tensor_bf_main_amd import workloads as wl

    log(f"generating workload 5: {args.sets} datasets")
    sets = wl.batch(n_sets=args.sets)
    dev = [{"data": ctx.to_device(s["data"]), "eps": s["eps"],
            "counts": ctx.empty((len(s["eps"]), s["data"].shape[0]), np.int64)} for s in sets]
    pairs = float(sum(s["data"].shape[0] ** 2 for s in sets))
    flops = float(sum(s["data"].shape[1] * s["data"].shape[0] * (s["data"].shape[0] + 1.0) for s in sets))
    log("plan: autotune + capture")
    t0 = time.perf_counter()
    plan = ctx.plan(dev, dtype=args.dtype)
    t_plan = time.perf_counter() - t0
    for _ in range(args.warmup):
        plan.run()
    st, t0 = None, time.perf_counter()
    for _ in range(args.steps):
        st = plan.run()
    t_graph = (time.perf_counter() - t0) / args.steps
    errors = pass_errors(ctx)  # the plan's last replay
    graph_counts = [dv["counts"].numpy() for dv in dev]
    log(f"graph: {t_graph * 1e3:.1f} ms/step ({pairs / t_graph:.3e} pairs/s)")
    # The same work as one call per dataset (eager, per-dataset launches + sync).
    for dv in dev[:8]:
        ctx.self_join_device(dv["data"], dv["eps"], dtype=args.dtype, counts=dv["counts"])
    t0 = time.perf_counter()
    for dv in dev:
        ctx.self_join_device(dv["data"], dv["eps"], dtype=args.dtype, counts=dv["counts"])
    t_each = time.perf_counter() - t0
    mismatches = sum(int((dv["counts"].numpy() != g).sum()) for dv, g in zip(dev, graph_counts))
    log(f"per-dataset launches: {t_each * 1e3:.1f} ms; graph speedup {t_each / t_graph:.2f}x")
    out = {"metric": f"pair-distances/s, {len(sets)} datasets x 4 eps in one persistent hipGraph launch",
           "value": pairs / t_graph, "unit": "pair-distances/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": t_graph * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": args.dtype, "data": "synthetic",
           "config": {"workload": f"batch_{len(sets)}", "sets": len(sets), "total_pairs": pairs,
                      "dims": sorted({int(s["data"].shape[1]) for s in sets})},
           "per_dataset_launches_ms": t_each * 1e3, "per_dataset_pairs_per_s": pairs / t_each,
           "graph_speedup_vs_per_dataset": t_each / t_graph, "plan_create_s": t_plan,
           "band_pairs": int(st["band_pairs"]), "fp64_refine_pairs": int(st["fp64_pairs"]),
           "candidates": int(st["candidates"]), "slices": int(st["slices"]),
           "stage_ms": {k: st[k] for k in ("ms_prep", "ms_join", "ms_refine", "ms_total")},
           "roofline": roofline(flops, st["ms_join"], args.dtype), "errors": errors, "tiles": plan.tiles(),
           "mismatches_graph_vs_eager": mismatches}
    if not args.skip_cpu:
        from oracle import oracle

        # Exact counts of a seeded few whole datasets against the oracle (among those of
        # at most 8192 rows, so that the check stays within seconds).
        small = [k for k, s in enumerate(sets) if s["data"].shape[0] <= 8192]
        pick = np.random.default_rng(7).choice(small, min(6, len(small)), replace=False) if small else []
        mism = 0
        for k in pick:
            oc, _ = oracle.join(sets[k]["data"], sets[k]["eps"])
            mism += int((graph_counts[k] != oc).sum())
        out["oracle_sets_checked"] = [int(k) for k in pick]
        out["oracle_mismatches"] = mism
    out["mismatches"] = mismatches + out.get("oracle_mismatches", 0)
    plan.close()
    for dv in dev:
        dv["data"].free()
        dv["counts"].free()
    return out


def copy_calibration(ctx, mib):
    """One hipMemcpy of a known size through the HIP runtime the library uses:
    the FETCH_SIZE / WRITE_SIZE counters it produces convert counter readings of
    the join into HBM bytes."""
    import tensor_bf_main_amd as bfj
    nbytes = mib << 20
    src = bfj.DeviceArray(ctx, (nbytes,), np.uint8)
    dst = bfj.DeviceArray(ctx, (nbytes,), np.uint8)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    rc = hip.hipMemcpy(dst.data_ptr(), src.data_ptr(), nbytes, 3)  # hipMemcpyDeviceToDevice
    rc = rc or hip.hipDeviceSynchronize()
    if rc:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    print(json.dumps({"copy_bytes": nbytes}), flush=True)
    src.free()
    dst.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=3, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--d", type=int, default=0, help="dimension of workload 2 (uniform; 0 = 128)")
    ap.add_argument("--sets", type=int, default=512)
    ap.add_argument("--dtype", default="", choices=["", "fp16", "bf16", "fp8", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0, help="query rows timed on the CPU (0 = auto, ~10-30 s)")
    ap.add_argument("--tile-rows", type=int, default=0, help="force the join's rows per wavefront (0 = library default)")
    ap.add_argument("--sub-steps", type=int, default=3, help="timed steps of each other configuration")
    ap.add_argument("--no-sub-configs", action="store_true",
                    help="headline only (by default workload 3 also measures workloads 1, 2, 4 and 5)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the join runs on a single MI355X (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import tensor_bf_main_amd as bfj
    from tensor_bf_main_amd import workloads as wl

    ctx = bfj.Context(0)
    if args.copy_calibration:
        copy_calibration(ctx, args.copy_calibration)
        ctx.close()
        return
    if args.workload == 5:
        args.dtype = args.dtype or "fp16"
        out = batch_workload(args, ctx)
    else:
        n = args.n or {1: 10_000, 2: 100_000, 3: 1_000_000, 4: 500_000}[args.workload]
        dtype = args.dtype or {1: "fp64", 2: "fp64", 3: "fp16", 4: "bf16"}[args.workload]
        if args.tile_rows:
            ctx.set_tile_rows(dtype, args.tile_rows)
        log(f"generating workload {args.workload}: n={n}")
        if args.d and args.workload != 2:
            sys.exit("--d applies to workload 2 only")
        x, eps = wl.config(args.workload, n=n, d=args.d or None)
        w2 = ("pair-distances/s at n=100k,d=128 (FP64-MFMA + FP64 band refine, exact counts)" if not args.d else
              f"pair-distances/s at n={n},d={args.d} ({dtype.upper()}-MFMA + FP64 band refine, exact counts)")
        names = {1: ("pair-distances/s at n=10k,d=32 (FP64-MFMA + FP64 band refine, exact counts)", "uniform_n10k_d32"),
                 2: (w2, "uniform" if not args.d else f"uniform_d{args.d}"),
                 3: ("pair-distances/s at n=1M,d=128 (FP16-MFMA + FP64 band refine, exact counts)",
                     "gmm_64c_sigma0.03"),
                 4: (f"pair-distances/s at n=500k,d=256 ({dtype.upper()}-MFMA + FP64 band refine, exact counts)",
                     "unit_norm_5pct_dups")}
        metric, wname = names[args.workload]
        out = single_dataset(args, ctx, x, eps, dtype, metric, wname)
        out.pop("fp64_counts", None)
        del x
        if args.workload == 3 and not args.n and not args.no_sub_configs:
            # Every other configuration at its stated size, measured in the same run.
            out["configs"] = sub_configs(args, ctx)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
