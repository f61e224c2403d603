"""Pins the CPU FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle fails it:
* exact-rational golden counts of the reference's own test files,
* bit-exact fma-chain distances against a rational emulation of
  t = q - c; acc = fma(t, t, acc) in ascending k (kernel_join.cu:184-189),
* the threshold T against sqrt(x) <= eps on neighbours of T (kernel_join.cu:194),
* closed-form lattices, eps = 0 with duplicates, eps >= diameter, symmetry,
  monotonicity, and a numpy cross-check,
* the reader against the reference's own dataset.cu compiled into oracle/_ref.
"""
import json
import math
import os
import tempfile
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fma_chain(q, c):
    """Exact emulation of acc = fma(t, t, acc), t = fl(q - c), ascending k."""
    acc = 0.0
    for a, b in zip(q, c):
        t = float(Fraction(a) - Fraction(b))  # correctly rounded subtraction
        acc = float(Fraction(t) * Fraction(t) + Fraction(acc))  # one rounding: fma
    return acc


def test_golden_reference_files():
    gold = json.load(open(os.path.join(GOLDEN, "expected_counts.json")))
    for name in ("test100.txt", "test128.txt"):
        rows = oracle.parse_text(open(os.path.join(GOLDEN, name)).read(), 18)
        g = gold[name]
        assert rows.shape == (g["n"], 18)
        counts, totals = oracle.join(rows, g["eps"])
        assert [int(t) for t in totals] == g["totals"]
        np.testing.assert_array_equal(counts, np.asarray(g["counts"]))
    # Independent numpy FP64 totals for the reference's two sample files.
    assert gold["test100.txt"]["totals"][gold["test100.txt"]["eps"].index(0.05)] == 544
    assert gold["test128.txt"]["totals"][gold["test128.txt"]["eps"].index(0.05)] == 828


def test_pair_d2_is_the_fma_chain_bit_for_bit():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((40, 23)) * np.exp(rng.uniform(-20, 20, (40, 1)))
    i = rng.integers(0, 40, 200)
    j = rng.integers(0, 40, 200)
    got = oracle.pair_d2(x, i, j)
    want = np.array([_fma_chain(x[a], x[b]) for a, b in zip(i, j)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # The same data summed without fma, or in another order, gives other bits:
    # the check above can tell those apart.
    nofma = np.array([sum_plain(x[a], x[b]) for a, b in zip(i, j)])
    assert not np.array_equal(nofma.view(np.uint64), want.view(np.uint64))


def sum_plain(q, c):
    acc = 0.0
    for a, b in zip(q, c):
        t = a - b
        acc = acc + t * t
    return acc


@pytest.mark.parametrize("eps", [0.0, 1e-300, 5e-324, 0.1, 1.0, 2.0, 1e10, 3.7e153, 1e200, math.pi])
def test_threshold_matches_sqrt_predicate(eps):
    T = oracle.threshold(eps)
    assert math.sqrt(T) <= eps
    up = math.nextafter(T, math.inf)  # inf past the largest double, without a warning
    assert math.isinf(up) or not (math.sqrt(up) <= eps)
    x = T
    for _ in range(50):
        x = math.nextafter(x, 0.0)
        assert math.sqrt(x) <= eps
    assert oracle.threshold(-1.0) < 0 and math.isinf(oracle.threshold(float("inf")))


def test_threshold_random():
    rng = np.random.default_rng(9)
    for eps in rng.uniform(0, 3, 2000):
        T = oracle.threshold(float(eps))
        assert math.sqrt(T) <= eps < math.sqrt(np.nextafter(T, np.inf)) or T == np.nextafter(T, np.inf)


def test_lattice_closed_form():
    # Points of {0..m-1}^2: neighbours within 1 are the 4-neighbourhood, within sqrt(2)
    # the 8-neighbourhood, within 2 additionally the points 2 apart on an axis.
    m = 9
    g = np.array([(a, b) for a in range(m) for b in range(m)], dtype=np.float64)
    counts, totals = oracle.join(g, [1.0, math.sqrt(2.0), 2.0])

    def expect(offsets):
        out = []
        for a, b in g.astype(int):
            out.append(sum(1 for da, db in offsets if 0 <= a + da < m and 0 <= b + db < m))
        return np.array(out)

    four = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    eight = four + [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    twelve = eight + [(2, 0), (-2, 0), (0, 2), (0, -2)]
    np.testing.assert_array_equal(counts[0], expect(four))
    np.testing.assert_array_equal(counts[1], expect(eight))
    np.testing.assert_array_equal(counts[2], expect(twelve))
    assert int(totals[0]) == m * m + 2 * 2 * m * (m - 1)


def test_duplicates_eps0_and_diameter():
    rng = np.random.default_rng(1)
    base = rng.random((50, 6))
    x = np.concatenate([base, base[:7], base[:3]])
    n = x.shape[0]
    counts, totals = oracle.join(x, [0.0, 10.0, -0.5])
    assert int(totals[0]) == n + 2 * (3 * 3 + 4)  # rows 0..2 three times (3 pairs each), 3..6 twice
    assert int(totals[1]) == n * n
    assert int(totals[2]) == 0 and counts[2].sum() == 0


def test_symmetry_monotonicity_rowsubset():
    rng = np.random.default_rng(2)
    x = rng.random((400, 9))
    eps = np.linspace(0.05, 0.9, 7)
    counts, totals = oracle.join(x, eps)
    assert np.all(np.diff(counts, axis=0) >= 0)
    assert all(int(counts[k].sum()) % 2 == 0 for k in range(len(eps)))
    # pairwise relation is symmetric
    d2 = oracle.pair_d2(x, np.repeat(np.arange(400), 400), np.tile(np.arange(400), 400)).reshape(400, 400)
    assert np.array_equal(d2, d2.T)
    rows = np.array([5, 77, 399, 0])
    np.testing.assert_array_equal(oracle.join_rows(x, rows, eps), counts[:, rows])
    c2, t2 = oracle.join(x, eps, 100, 250)
    np.testing.assert_array_equal(c2, counts[:, 100:250])
    assert np.all(t2 == c2.sum(axis=1) + 150)


def test_numpy_cross_check_away_from_boundary():
    rng = np.random.default_rng(4)
    x = rng.random((300, 7))
    diff = x[:, None, :] - x[None, :, :]
    d2 = np.einsum("ijk,ijk->ij", diff, diff)
    for eps in (0.2, 0.4, 0.6):
        margin = np.abs(d2 - eps * eps).min()
        assert margin > 1e-12
        counts, totals = oracle.join(x, eps)
        want = (d2 <= eps * eps).sum(axis=1) - 1
        np.testing.assert_array_equal(counts[0], want)


def test_planted_off_by_one_threshold_is_caught():
    """An oracle that used T = nextafter(T, -inf) (strict inequality at the boundary)
    differs from the real one exactly on boundary pairs; the boundary check sees it."""
    rng = np.random.default_rng(6)
    x = rng.random((60, 5))
    d2 = oracle.pair_d2(x, [3], [17])[0]
    eps = math.sqrt(d2)
    good, _ = oracle.join(x, eps)
    assert good[0][3] >= 1 and good[0][17] >= 1
    planted_T = np.nextafter(oracle.threshold(eps), -np.inf)
    planted = np.array([[sum(1 for j in range(60) if j != i and oracle.pair_d2(x, [i], [j])[0] <= planted_T)
                         for i in range(60)]])
    assert not np.array_equal(planted, good)


def test_parse_text_rules():
    txt = "1, 2, 3\n4 5\n6,7,8, 9 10"
    assert oracle.parse_text(txt, 3).tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 9]]
    assert oracle.parse_text("1 2 x 3 4", 2).tolist() == [[1, 2]]
    assert oracle.parse_text("", 2).shape == (0, 2)


def _ref_available():
    return oracle.ref_import is not None and os.path.exists(oracle._REF_IMPORT)


@pytest.mark.skipif(not os.path.exists(oracle._REF_IMPORT), reason="reference reader not built (oracle/_ref)")
def test_parse_text_matches_reference_reader():
    rng = np.random.default_rng(8)
    vals = rng.random(90 * 3 + 17)
    cases = [
        ", ".join(f"{v:.6f}" for v in vals),
        "\n".join(" ".join(f"{v:.9g}" for v in vals[k:k + 11]) for k in range(0, vals.size, 11)),
        ",".join(f"{v:.3e}" for v in vals),
        " , ".join(f"{v}" for v in vals[:200]),
    ]
    for txt in cases:
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(txt)
            path = f.name
        try:
            ref = oracle.ref_import(path)
            mine = oracle.parse_text(txt, 90)
            assert ref.shape == mine.shape
            np.testing.assert_array_equal(ref, mine)
        finally:
            os.unlink(path)


def test_thread_count_and_counts_independent_of_threads():
    """num_threads reports OpenMP's thread count (OMP_NUM_THREADS in a fresh process),
    and the counts do not depend on how many threads share the rows."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", "from oracle import oracle; print(oracle.num_threads())"],
                       cwd=root, env=dict(os.environ, OMP_NUM_THREADS="3"), capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "3", r.stderr[-2000:]
    assert oracle.num_threads() >= 1
    rng = np.random.default_rng(11)
    x = rng.random((700, 9))
    eps = [0.35, 0.5]
    c1, t1 = oracle.join(x, eps, nthreads=1)
    c5, t5 = oracle.join(x, eps, nthreads=5)
    assert np.array_equal(c1, c5) and np.array_equal(t1, t5)
    rows = rng.choice(700, size=50, replace=False)
    assert np.array_equal(oracle.join_rows(x, rows, eps, nthreads=1), oracle.join_rows(x, rows, eps, nthreads=7))
    assert np.array_equal(oracle.join_rows(x, rows, eps, nthreads=3), c1[:, rows])
