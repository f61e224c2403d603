"""CPU FP64 oracle for the brute-force epsilon self-join -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
import this module, and only as the checker.  The product package
(``tensor_bf_main_amd``) never imports it.

Functions and the reference code they restate (jeevana-swaroop/tensor_bf_main):

* :func:`threshold` -- the predicate ``sqrt(d2) <= eps`` (kernel_join.cu:194) as
  ``d2 <= T``; C implementation in bfj_oracle.c.
* :func:`pair_d2` -- direct-form squared distance, ascending k, ``t = q-c;
  acc = fma(t, t, acc)`` (kernel_join.cu:184-189).
* :func:`join` / :func:`join_rows` -- per-point neighbour counts (self
  excluded) and the reference's ordered, self-inclusive total
  (gpu_join.cu:318-321, kernel_join.cu:184).
* :func:`parse_text` -- the text reader, dataset.cu:24-38: values separated by
  ``", "`` or whitespace, grouped into rows of ``dim`` values regardless of line
  breaks, a trailing partial row dropped.
* :func:`reference_total` -- the reference's own arithmetic of search modes 1-4
  (kernel_join.cu:168-524), for the bug-for-bug totals.

How each is pinned (tests/test_oracle.py, tests/test_reference_modes.py): golden
vectors of the reference's own data files, exact-rational emulations, closed forms
and invariants, and the reference's own reader compiled from its sources into _ref/.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BFJ_ORACLE_LIB selects another build of bfj_oracle.c (the sanitizer build in tests).
_LIB_PATH = os.environ.get("BFJ_ORACLE_LIB") or os.path.join(_HERE, "libbfj_oracle.so")
_REF_IMPORT = os.path.join(_HERE, "_ref", "ref_import")

_lib = None


def build(verbose: bool = False) -> None:
    """Compile libbfj_oracle.so (and _ref/ when the reference tree is present)."""
    out = None if verbose else subprocess.DEVNULL
    subprocess.check_call(["make", "-C", _HERE, "libbfj_oracle.so"], stdout=out)
    if os.path.isdir("/root/reference"):
        subprocess.call(["make", "-C", _HERE, "ref"], stdout=out, stderr=out)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64p = ctypes.POINTER(ctypes.c_int64)
    f64p = ctypes.POINTER(ctypes.c_double)
    lib.bfo_threshold.restype = ctypes.c_double
    lib.bfo_threshold.argtypes = [ctypes.c_double]
    lib.bfo_pair_d2.restype = ctypes.c_double
    lib.bfo_pair_d2.argtypes = [f64p, f64p, ctypes.c_int32]
    lib.bfo_pair_d2_batch.restype = ctypes.c_int
    lib.bfo_pair_d2_batch.argtypes = [f64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, i64p, i64p,
                                      ctypes.c_int64, f64p]
    lib.bfo_join_rows.restype = ctypes.c_int
    lib.bfo_join_rows.argtypes = [f64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, i64p, ctypes.c_int64,
                                  f64p, ctypes.c_int32, i64p, ctypes.c_int32]
    lib.bfo_join.restype = ctypes.c_int
    lib.bfo_join.argtypes = [f64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, f64p, ctypes.c_int32,
                             ctypes.c_int64, ctypes.c_int64, i64p, ctypes.POINTER(ctypes.c_uint64),
                             ctypes.c_int32]
    lib.bfo_num_threads.restype = ctypes.c_int
    lib.bfo_round_half.restype = ctypes.c_double
    lib.bfo_round_half.argtypes = [ctypes.c_double]
    lib.bfo_reference_total.restype = ctypes.c_int
    lib.bfo_reference_total.argtypes = [f64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    _lib = lib
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _as_data(data):
    data = np.ascontiguousarray(data, dtype=np.float64)
    if data.ndim != 2:
        raise ValueError("data must be 2-D (n, d)")
    return data


def num_threads() -> int:
    return int(_load().bfo_num_threads())


def threshold(eps: float) -> float:
    """Largest double T with sqrt(T) <= eps."""
    return float(_load().bfo_threshold(float(eps)))


def pair_d2(data, i, j) -> np.ndarray:
    """Direct-form FP64 squared distances for index arrays i, j."""
    data = _as_data(data)
    pi = np.ascontiguousarray(np.atleast_1d(i), dtype=np.int64)
    pj = np.ascontiguousarray(np.atleast_1d(j), dtype=np.int64)
    out = np.empty(pi.size, dtype=np.float64)
    n, d = data.shape
    rc = _load().bfo_pair_d2_batch(_ptr(data, ctypes.c_double), n, d, d, _ptr(pi, ctypes.c_int64),
                                   _ptr(pj, ctypes.c_int64), pi.size, _ptr(out, ctypes.c_double))
    if rc:
        raise ValueError(f"bfo_pair_d2_batch failed ({rc})")
    return out


def join_rows(data, rows, eps, nthreads: int = 0) -> np.ndarray:
    """Counts (self excluded) of ``rows`` against all points: shape (n_eps, len(rows))."""
    data = _as_data(data)
    eps = np.ascontiguousarray(np.atleast_1d(eps), dtype=np.float64)
    rows = np.ascontiguousarray(np.atleast_1d(rows), dtype=np.int64)
    n, d = data.shape
    counts = np.zeros((eps.size, rows.size), dtype=np.int64)
    rc = _load().bfo_join_rows(_ptr(data, ctypes.c_double), n, d, d, _ptr(rows, ctypes.c_int64), rows.size,
                               _ptr(eps, ctypes.c_double), eps.size, _ptr(counts, ctypes.c_int64), nthreads)
    if rc:
        raise ValueError(f"bfo_join_rows failed ({rc})")
    return counts


def join(data, eps, row_begin: int = 0, row_end: int | None = None, nthreads: int = 0):
    """Full join.  Returns (counts[n_eps, rows] self excluded, totals[n_eps] self included)."""
    data = _as_data(data)
    eps = np.ascontiguousarray(np.atleast_1d(eps), dtype=np.float64)
    n, d = data.shape
    if row_end is None:
        row_end = n
    counts = np.zeros((eps.size, max(row_end - row_begin, 0)), dtype=np.int64)
    totals = np.zeros(eps.size, dtype=np.uint64)
    rc = _load().bfo_join(_ptr(data, ctypes.c_double), n, d, d, _ptr(eps, ctypes.c_double), eps.size,
                          row_begin, row_end, _ptr(counts, ctypes.c_int64), _ptr(totals, ctypes.c_uint64),
                          nthreads)
    if rc:
        raise ValueError(f"bfo_join failed ({rc})")
    return counts, totals


def round_half(x: float) -> float:
    """binary16 rounding (nearest even) of a double, as the reference-arithmetic
    restatement uses it."""
    return float(_load().bfo_round_half(float(x)))


def reference_total(data, eps: float, mode: int, compute_prec: int = 16, accum_prec: int = 32) -> int:
    """The reference's own total for search mode 1-4 with its COMPUTE_PREC / ACCUM_PREC
    arithmetic (kernel_join.cu:168-524 restated in bfj_oracle.c, bfo_reference_total):
    ordered pairs, self included."""
    data = _as_data(data)
    n, d = data.shape
    tot = ctypes.c_uint64(0)
    rc = _load().bfo_reference_total(_ptr(data, ctypes.c_double), n, d, d, float(eps), int(mode), int(compute_prec),
                                     int(accum_prec), ctypes.byref(tot))
    if rc:
        raise ValueError(f"bfo_reference_total failed ({rc})")
    return int(tot.value)


_NUMBER = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.IGNORECASE)
_COMMA_WS = re.compile(r",\s*")


def parse_text(text: str, dim: int) -> np.ndarray:
    """dataset.cu:24-38 restated.  Each loop step is ``fscanf("%lf, ")``: skip blanks, read one
    number, then consume a comma and the blanks after it when a comma comes next.  Values are
    grouped into rows of ``dim`` regardless of line breaks, a trailing partial row is dropped,
    and reading stops at the first thing that is not a number (both fscanf calls return 0)."""
    vals = []
    pos = 0
    while True:
        m = _NUMBER.match(text, pos)
        if not m:
            break
        vals.append(float(m.group(1)))
        pos = m.end()
        c = _COMMA_WS.match(text, pos)
        if c:
            pos = c.end()
    m = len(vals) // dim
    return np.asarray(vals[: m * dim], dtype=np.float64).reshape(m, dim)


def ref_import(path: str):
    """Rows parsed by the reference's own dataset.cu (INPUT_DATA_DIM=90), or None if not built."""
    if not os.path.exists(_REF_IMPORT):
        return None
    out = subprocess.run([_REF_IMPORT, path], capture_output=True, text=True, check=True).stdout.split("\n")
    m, dim = (int(x) for x in out[0].split())
    rows = [list(map(float, ln.split())) for ln in out[1:1 + m]]
    return np.asarray(rows, dtype=np.float64).reshape(m, dim)
