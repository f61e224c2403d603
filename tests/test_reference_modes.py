"""The reference's own arithmetic for search modes 1-4 (bug-for-bug totals).

CPU tests pin the oracle's restatement (oracle/bfj_oracle.c, bfo_reference_total):
binary16 rounding against numpy, the FP64 case against the exact FP64 oracle, an
independent numpy restatement for the cases where every operation is a single IEEE
rounding numpy can express, and an integer lattice on which every mode is exact.
GPU tests compare the library (bfj_reference_mode_join) with the oracle.

Modes 1 and 2 (kernel_join.cu:168-239) are fully specified IEEE arithmetic.  Modes 3
and 4 (kernel_join.cu:245-524) ran on tensor cores whose summation order is not
documented; for them the column-order FP32 sum is the library's definition, so their
parity with the reference itself is unpinned.
"""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle
from tensor_bf_main_amd import workloads as wl

COMBOS = [(1, 16, 32), (1, 16, 64), (1, 32, 32), (1, 32, 64), (1, 64, 32), (1, 64, 64),
          (2, 16, 32), (2, 16, 64), (2, 32, 32), (2, 32, 64), (2, 64, 32), (2, 64, 64),
          (3, 16, 32), (4, 16, 32)]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _eps(x, q):
    n = x.shape[0]
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    return float(np.quantile(d[np.triu_indices(n, 1)], q))


# ---- numpy restatement (independent of the C oracle) ---------------------------------

def _np_total(x, eps, mode, cp, ap):
    """All pairs at once with numpy's IEEE binary16/32/64 operations (each a single
    rounding).  Only for the combinations that need no fused multiply-add."""
    ct = {16: np.float16, 32: np.float32, 64: np.float64}[cp]
    at = {32: np.float32, 64: np.float64}[ap]
    n, d = x.shape
    dpad = (d + 15) // 16 * 16
    X = np.zeros((n, dpad), dtype=ct)
    X[:, :d] = x.astype(ct)
    acc = np.zeros((n, n), dtype=at)
    if mode in (1, 2):
        for k in range(d):
            q = X[:, k][:, None]
            c = X[:, k][None, :]
            if mode == 1:
                t = (q - c).astype(ct)
                acc = acc + (t * t).astype(ct).astype(at)
            else:
                qm2 = (ct(-2.0) * q).astype(ct)
                q2 = q.astype(at) * q.astype(at)
                acc = acc + (((qm2 * c).astype(ct).astype(at) + q2) + (c * c).astype(ct).astype(at))
    else:
        X32 = X.astype(np.float32)
        cn = np.zeros((n, dpad // 16), dtype=np.float32)
        for ch in range(dpad // 16):
            for r in range(16):
                v = X32[:, ch * 16 + r]
                cn[:, ch] = cn[:, ch] + v * v
        for ch in range(dpad // 16):
            s = None
            for r in range(16):
                k = ch * 16 + r
                if mode == 3:
                    dd = (X[:, k][:, None] - X[:, k][None, :]).astype(np.float16).astype(np.float32)
                    p = dd * dd
                else:
                    p = (np.float16(-2.0) * X[:, k]).astype(np.float16).astype(np.float32)[:, None] * X32[:, k][None, :]
                s = p if s is None else s + p
            if mode == 3:
                acc = acc + s
            else:
                acc = (acc + (cn[:, ch][None, :] + s)) + cn[:, ch][:, None]
    v = np.abs(acc) if mode in (2, 4) else acc
    with np.errstate(invalid="ignore"):
        return int((np.sqrt(v) <= at(eps)).sum())


def test_round_half_matches_numpy():
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 6, 20000),
        # ties and near-ties of the binary16 grid, subnormals, the overflow edge
        (np.arange(-2000, 2000) + 0.5) * 2.0 ** -10,
        (np.arange(0, 3000) + 0.5) * 2.0 ** -24,
        np.array([65504.0, 65519.99, 65520.0, -65520.0, 2.0 ** -25, 2.0 ** -25 * 1.0000001, 0.0, -0.0]),
    ])
    ours = np.array([oracle.round_half(v) for v in vals])
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).astype(np.float64)
    np.testing.assert_array_equal(ours, ref)


@pytest.mark.parametrize("n,d,q", [(200, 18, 0.01), (150, 90, 0.05), (97, 7, 0.2)])
def test_fp64_mode1_is_the_exact_oracle(n, d, q):
    x = wl.uniform(n, d, 11 + n)
    eps = _eps(x, q)
    _, tot = oracle.join(x, eps)
    assert oracle.reference_total(x, eps, 1, 64, 64) == int(tot[0])


@pytest.mark.parametrize("mode,cp,ap", [c for c in COMBOS if c[1] != c[2]])
@pytest.mark.parametrize("seed", [1, 2])
def test_restatement_matches_numpy(mode, cp, ap, seed):
    x = wl.gmm(120, 21, 4, 0.05, seed) if seed == 1 else wl.uniform(110, 18, seed) * 3.0 - 1.0
    for q in (0.01, 0.1):
        eps = _eps(x, q)
        assert oracle.reference_total(x, eps, mode, cp, ap) == _np_total(x, eps, mode, cp, ap), (q, eps)


@pytest.mark.parametrize("mode,cp,ap", COMBOS)
def test_integer_lattice_is_exact(mode, cp, ap):
    """Small integer coordinates: every product and sum is exact in every type, so each
    mode returns the exact total."""
    rng = np.random.default_rng(mode * 100 + cp + ap)
    x = rng.integers(-4, 5, size=(90, 6)).astype(np.float64)
    x[10:20] = x[0]  # duplicates
    for eps in (0.0, 1.0, np.sqrt(5.5), 4.5):
        _, tot = oracle.join(x, eps)
        assert oracle.reference_total(x, eps, mode, cp, ap) == int(tot[0]), eps


# ---- exact-rational emulation of the same-type combinations ---------------------------
# When COMPUTE_PREC == ACCUM_PREC nvcc's default contraction (-fmad=true) fuses a product
# into the sum that consumes it.  numpy has no fused multiply-add, so these three
# combinations are pinned by an emulation in exact rationals with one correctly rounded
# step per operation, of the reference's own expressions (kernel_join.cu:184-189 and
# :222-231), in both the fused and the unfused reading; the data are chosen so that the
# two readings give different totals, so the check can tell them apart.

def _round(x: Fraction, p: int, emin: int) -> Fraction:
    """x rounded to nearest even in binary floating point with p-bit significands and
    minimum exponent emin (binary32: 24, -126; binary64: 53, -1022)."""
    if x == 0:
        return Fraction(0)
    sgn, a = (-1 if x < 0 else 1), abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, emin)
    ulp = Fraction(2) ** (e - p + 1)
    m = a / ulp
    q, r = divmod(m.numerator, m.denominator)
    if 2 * r > m.denominator or (2 * r == m.denominator and q % 2 == 1):
        q += 1
    return sgn * q * ulp


def _emulate_d2(x, mode, prec, fused):
    """acc for every ordered pair of rows of x (compute = accumulate = binary32 or 64)."""
    p, emin = (24, -126) if prec == 32 else (53, -1022)
    R = lambda v: _round(v, p, emin)  # noqa: E731
    X = [[R(Fraction(float(v))) for v in row] for row in x]  # convertDataset to COMPUTE_TYPE
    n, d = len(X), len(X[0])
    out = np.zeros((n, n))
    for i in range(n):
        q = X[i]
        if mode == 2:
            pt = [R(-2 * v) for v in q]   # (COMPUTE_TYPE)(-2.0) * queryCoord
            q2 = [R(v * v) for v in q]     # (ACCUM_TYPE)q * (ACCUM_TYPE)q, its own statement
        for j in range(n):
            c = X[j]
            acc = Fraction(0)
            for k in range(d):
                if mode == 1:                       # acc += (q - c) * (q - c)
                    t = R(q[k] - c[k])
                    acc = R(acc + t * t) if fused else R(acc + R(t * t))
                else:                               # acc += (pt * c) + q2 + c2, c2 = c * c
                    if fused:
                        s_ = R(c[k] * c[k] + R(pt[k] * c[k] + q2[k]))
                    else:
                        s_ = R(R(R(pt[k] * c[k]) + q2[k]) + R(c[k] * c[k]))
                    acc = R(acc + s_)
            out[i, j] = float(acc)
    return out


def _hits(acc, eps, prec, mode):
    v = np.abs(acc) if mode == 2 else acc
    if prec == 32:
        return int((np.sqrt(v.astype(np.float32)) <= np.float32(eps)).sum())
    return int((np.sqrt(v) <= eps).sum())


@pytest.mark.parametrize("mode,prec", [(1, 32), (2, 32), (2, 64)])
def test_same_type_contraction_pinned_by_exact_emulation(mode, prec):
    rng = np.random.default_rng(40 + mode + prec)
    x = rng.uniform(0.3, 0.45, size=(34, 19)) * (1.0 + rng.uniform(-1e-3, 1e-3, size=(34, 1)))
    fused = _emulate_d2(x, mode, prec, True)
    plain = _emulate_d2(x, mode, prec, False)
    ft = (lambda v: np.sqrt(np.abs(v).astype(np.float32))) if prec == 32 else (lambda v: np.sqrt(np.abs(v)))
    sf, sp = ft(fused), ft(plain)
    # eps at the square roots where the two readings disagree: one counts the pair, the other not
    split = np.argwhere(sf != sp)
    assert split.shape[0] >= 4, "the data do not separate the fused and unfused readings"
    eps_list = sorted({float(min(sf[i, j], sp[i, j])) for i, j in split[:40]})
    differ = 0
    for eps in eps_list:
        want = _hits(fused, eps, prec, mode)
        differ += want != _hits(plain, eps, prec, mode)
        assert oracle.reference_total(x, eps, mode, prec, prec) == want, eps
    assert differ >= 1, "no eps separates the readings: the check could not fail"


def test_unsupported_combinations_raise():
    x = np.zeros((4, 3))
    for mode, cp, ap in [(3, 32, 32), (4, 64, 64), (1, 16, 16), (5, 16, 32)]:
        with pytest.raises(ValueError):
            oracle.reference_total(x, 1.0, mode, cp, ap)


# ---- GPU: the library against the oracle ---------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("mode,cp,ap", COMBOS)
def test_library_matches_oracle(ctx, mode, cp, ap):
    cases = [(wl.gmm(700, 90, 6, 0.04, 5), (0.0, 0.005, 0.05)),   # the reference's INPUT_DATA_DIM
             (wl.uniform(333, 18, 8), (0.01, 0.3)),               # tail tiles, d not a multiple of 16
             (wl.uniform(9, 33, 9), (0.5,))]                      # fewer rows than a tile
    for x, qs in cases:
        for q in qs:
            eps = _eps(x, q) if q > 0 else 0.0
            assert ctx.reference_mode_join(x, eps, mode, cp, ap) == oracle.reference_total(x, eps, mode, cp, ap), \
                (x.shape, q)
    assert ctx.reference_mode_join(np.zeros((0, 5)), 1.0, mode, cp, ap) == 0
    assert ctx.reference_mode_join(wl.uniform(20, 4, 1), -1.0, mode, cp, ap) == 0


@pytest.mark.gpu
def test_library_rejects_unsupported(ctx):
    import tensor_bf_main_amd as bfj
    x = wl.uniform(10, 4, 1)
    for mode, cp, ap in [(3, 32, 32), (4, 64, 64), (1, 16, 16), (0, 16, 32), (5, 16, 32)]:
        with pytest.raises(bfj.JoinError):
            ctx.reference_mode_join(x, 0.5, mode, cp, ap)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_cli_reference_arithmetic(tmp_path, mode):
    """bfjoin_main --reference-arithmetic prints the reference's own total (default
    COMPUTE_PREC 16 / ACCUM_PREC 32) on the reference's test file."""
    import tensor_bf_main_amd as bfj
    rows = oracle.parse_text(open(os.path.join(GOLDEN, "test128.txt")).read(), 18)
    want = oracle.reference_total(rows, 0.05, mode, 16, 32)
    r = subprocess.run([bfj.CLI_PATH, os.path.join(GOLDEN, "test128.txt"), "0.05", str(mode), "0", "--dim", "18",
                        "--reference-arithmetic", "--log", str(tmp_path / "log.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Total result set size: {want}" in r.stdout
    assert ", reference," in open(tmp_path / "log.csv").read()
