"""Seeded synthetic workloads for the epsilon self-join.

The reference ships no generator; its runs read text files (main.cu:30-36).
These generators produce the five workload shapes the join is measured on:

1. ``uniform``: n=10,000, d=32, U[0,1) (CPU-runnable).
2. ``uniform``: n=100,000, d=128, eps for ~64 neighbours (FP64-MFMA join).
3. ``gmm``: n=1,000,000, d=128, 64 centres U[0,1)^128, sigma=0.03, eps for ~100 neighbours.
4. ``unit_dups``: n=500,000, d=256 unit-norm rows, 5% near-duplicates, eps for ~32 neighbours.
5. ``batch``: 512 datasets, n log-uniform in [1000, 65536], d in {16, 64, 128, 200}, 4 eps each.

eps is an order statistic of the FP64 pair distances of a seeded subsample, so it
is the distance of a real pair and the boundary case is exercised.
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 1000


def uniform(n: int, d: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, d))


def gmm(n: int, d: int, centers: int, sigma: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = rng.random((centers, d))
    lab = rng.integers(0, centers, size=n)
    x = c[lab]
    x += rng.standard_normal((n, d)) * sigma
    return x


def unit_dups(n: int, d: int, seed: int, dup_frac: float = 0.05, jitter: float = 1e-4) -> np.ndarray:
    """Standard-normal rows, L2-normalised; a fraction of rows are replaced by
    copies of other rows plus N(0, jitter^2) noise."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    m = int(round(dup_frac * n))
    if m:
        dst = rng.choice(n, size=m, replace=False)
        src = rng.integers(0, n, size=m)
        x[dst] = x[src] + rng.standard_normal((m, d)) * jitter
    return x


def pair_distances(x: np.ndarray) -> np.ndarray:
    """All unordered FP64 pair distances of a (small) sample, direct differences."""
    m = x.shape[0]
    out = []
    for i in range(m - 1):
        diff = x[i + 1:] - x[i]
        out.append(np.sqrt(np.einsum("ij,ij->i", diff, diff)))
    return np.concatenate(out) if out else np.zeros(0)


def eps_quantile(x: np.ndarray, q, sample: int, seed: int):
    """Order statistic floor(q * (M - 1)) of the sorted pair distances of a seeded
    subsample (without replacement) of ``sample`` rows.  ``q`` may be a list."""
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    idx = rng.choice(n, size=min(sample, n), replace=False)
    dist = np.sort(pair_distances(x[idx]))
    qs = np.atleast_1d(np.asarray(q, dtype=np.float64))
    picks = [float(dist[min(int(np.floor(qq * (dist.size - 1))), dist.size - 1)]) for qq in qs]
    return picks if np.ndim(q) else picks[0]


def config(k: int, n: int | None = None, d: int | None = None):
    """(data, eps) of workload k (1..4) with optional reduced n (same recipe, same seed);
    d overrides the dimension of the uniform workload 2."""
    seed = SEED_BASE + k
    if k == 1:
        n = n or 10_000
        x = uniform(n, 32, seed)
        return x, eps_quantile(x, 0.005, 1000, seed + 100)
    if k == 2:
        n = n or 100_000
        x = uniform(n, d or 128, seed)
        return x, eps_quantile(x, 64.0 / (n - 1), 2000, seed + 100)
    if k == 3:
        n = n or 1_000_000
        x = gmm(n, 128, 64, 0.03, seed)
        return x, eps_quantile(x, 100.0 / (n - 1), 2000, seed + 100)
    if k == 4:
        n = n or 500_000
        x = unit_dups(n, 256, seed)
        return x, eps_quantile(x, 32.0 / (n - 1), 2000, seed + 100)
    raise ValueError("workloads 1..4 are single datasets; use batch() for 5")


def batch(n_sets: int = 512, seed: int = SEED_BASE + 5, n_min: int = 1000, n_max: int = 65536,
          dims=(16, 64, 128, 200), quantiles=(0.001, 0.005, 0.01, 0.02)):
    """Workload 5: list of {"data", "eps"} dicts.  Each dataset's shape, kind and seed
    are drawn in order from one generator; the datasets themselves come from their own
    seeds, so they are generated in parallel (threads: numpy's generators release the
    GIL) with the same values as one after the other."""
    from concurrent.futures import ThreadPoolExecutor
    import os

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_sets):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        d = int(rng.choice(dims))
        sub = int(rng.integers(0, 2**31))
        specs.append((n, d, sub, rng.random() < 0.5))

    def make(spec):
        n, d, sub, is_uniform = spec
        x = uniform(n, d, sub) if is_uniform else gmm(n, d, 8, 0.05, sub)
        return {"data": x, "eps": eps_quantile(x, list(quantiles), 500, sub + 1)}

    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        return list(ex.map(make, specs))
