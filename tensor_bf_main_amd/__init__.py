"""MI355X brute-force Euclidean epsilon self-join.

Python plumbing over the C ABI in ``include/bfjoin.h`` (``libbfjoin.so``, built
from ``csrc/``).  The join itself runs in hand-written gfx950 kernels; this module
only moves pointers.  There is no CPU fallback: if the library is missing or no
GPU is present, calls raise.

Reference semantics (jeevana-swaroop/tensor_bf_main): for every ordered pair of
rows (i, j) the pair is a hit when ``sqrt(d2) <= eps``, with ``d2`` the FP64 direct
form of kernel_join.cu:184-189; ``GPUJoinMainBruteForce`` (gpu_join.cu:210-322)
returns the self-inclusive total.  :meth:`Context.self_join` returns that total
per eps and per-point counts with self excluded.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbfjoin.so")
CLI_PATH = os.path.join(_HERE, "bfjoin_main")

DTYPES = {"fp64": 0, "fp16": 1, "bf16": 2, "fp8": 3}
DTYPE_NAMES = {v: k for k, v in DTYPES.items()}
MAX_EPS = 8
MAX_DIM = 256

BFJ_OK = 0
BFJ_ERR_NOMEM = 3
BFJ_ERR_CAPACITY = 6
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "DEVICE", 3: "NOMEM", 4: "UNSUPPORTED", 5: "INTERNAL", 6: "CAPACITY"}


class JoinError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bfjoin {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_uint64),
        ("band_pairs", ctypes.c_uint64),
        ("band_capacity", ctypes.c_uint64),
        ("band_retries", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("k_padded", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("max_band_halfwidth", ctypes.c_double),
        ("ms_prep", ctypes.c_double),
        ("ms_join", ctypes.c_double),
        ("ms_refine", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("fp64_pairs", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64),
        ("slices", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Errors(ctypes.Structure):
    """bfj_errors: accuracy of a pass's approximate squared distances."""
    _fields_ = [
        ("cand_total", ctypes.c_uint64),
        ("cand_pairs", ctypes.c_uint64),
        ("tile_pairs", ctypes.c_uint64),
        ("violations", ctypes.c_uint64),
        ("cand_max_abs_err", ctypes.c_double),
        ("tile_max_abs_err", ctypes.c_double),
        ("max_acc_ratio", ctypes.c_double),
        ("max_quant_ratio", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Dataset(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("d", ctypes.c_int32),
        ("ld", ctypes.c_int64),
        ("n_eps", ctypes.c_int32),
        ("eps", ctypes.POINTER(ctypes.c_double)),
        ("counts", ctypes.c_void_p),
        ("totals", ctypes.POINTER(ctypes.c_uint64)),
    ]


_lib = None
ABI_VERSION = 5  # BFJ_ABI_VERSION of include/bfjoin.h that Stats and the prototypes mirror
CHUNK_RECORDS = 4096  # BFJ_CHUNK_RECORDS: candidate records (16 bytes each) per stream chunk


def load_library(path: str = LIB_PATH):
    """Load libbfjoin.so.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (or make -C csrc)")
    lib = ctypes.CDLL(path)
    vp, i32, i64, f64p = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    sp = ctypes.POINTER(Stats)
    sig = {
        "bfj_last_error": (ctypes.c_char_p, []),
        "bfj_abi_version": (ctypes.c_int, []),
        "bfj_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
        "bfj_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "bfj_destroy": (ctypes.c_int, [vp]),
        "bfj_self_join": (ctypes.c_int, [vp, vp, i64, i32, i64, f64p, i32, i32, i64, i64, vp, u64p, sp]),
        "bfj_self_join_device": (ctypes.c_int, [vp, vp, i64, i32, i64, f64p, i32, i32, i64, i64, vp, u64p, sp]),
        "bfj_self_join_device_noalloc": (ctypes.c_int, [vp, vp, i64, i32, i64, f64p, i32, i32, i64, i64, vp, u64p,
                                                        sp]),
        "bfj_self_join_batch": (ctypes.c_int, [vp, ctypes.POINTER(Dataset), i32, i32, i32, sp]),
        "bfj_self_join_pairs": (ctypes.c_int, [vp, vp, i64, i32, i64, ctypes.c_double, i32, vp, ctypes.c_uint64,
                                               u64p, u64p, sp]),
        "bfj_reference_mode_join": (ctypes.c_int, [vp, vp, i64, i32, i64, ctypes.c_double, i32, i32, i32, u64p]),
        "bfj_plan_create": (ctypes.c_int, [vp, ctypes.POINTER(Dataset), i32, i32, ctypes.POINTER(vp)]),
        "bfj_plan_run": (ctypes.c_int, [vp, sp]),
        "bfj_plan_destroy": (ctypes.c_int, [vp]),
        "bfj_plan_tiles": (ctypes.c_int, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                          ctypes.POINTER(i32), i32, ctypes.POINTER(i32)]),
        "bfj_autotune": (ctypes.c_int, [vp, i32, i32, i64, ctypes.POINTER(i32)]),
        "bfj_set_tile_rows": (ctypes.c_int, [vp, i32, i32]),
        "bfj_set_reorder": (ctypes.c_int, [vp, i32]),
        "bfj_set_capacity": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint64]),
        "bfj_set_stream_limit": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "bfj_fail_allocation": (ctypes.c_int, [i64]),
        "bfj_pass_errors": (ctypes.c_int, [vp, ctypes.c_uint64, i64, ctypes.c_uint64, ctypes.POINTER(Errors)]),
        "bfj_tile_rows_options": (ctypes.c_int, [i32, i32, ctypes.POINTER(i32), i32, ctypes.POINTER(i32)]),
        "bfj_buffer_alloc": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(vp)]),
        "bfj_buffer_free": (ctypes.c_int, [vp, vp]),
        "bfj_buffer_upload": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64]),
        "bfj_buffer_download": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.bfj_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has C ABI version {lib.bfj_abi_version()}; this module is written for "
                           f"{ABI_VERSION}: rebuild the library")
    _lib = lib
    return lib


def _check(rc: int):
    if rc != BFJ_OK:
        raise JoinError(rc, _lib.bfj_last_error().decode())


def _dtype_id(dtype) -> int:
    if isinstance(dtype, str):
        if dtype not in DTYPES:
            raise ValueError(f"unknown dtype {dtype!r}; choose from {sorted(DTYPES)}")
        return DTYPES[dtype]
    return int(dtype)


def _eps_array(eps) -> np.ndarray:
    e = np.ascontiguousarray(np.atleast_1d(np.asarray(eps, dtype=np.float64)))
    if e.ndim != 1:
        raise ValueError("eps must be a scalar or a 1-D sequence")
    return e


def device_count() -> int:
    lib = load_library()
    n = ctypes.c_int(0)
    _check(lib.bfj_device_count(ctypes.byref(n)))
    return n.value


def tile_rows_options(dtype, d: int) -> list[int]:
    lib = load_library()
    opts = (ctypes.c_int32 * 8)()
    k = ctypes.c_int32(0)
    _check(lib.bfj_tile_rows_options(_dtype_id(dtype), d, opts, 8, ctypes.byref(k)))
    return [opts[i] for i in range(min(k.value, 8))]


class DeviceArray:
    """A device buffer owned by a :class:`Context` (row-major, C order)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = ctypes.c_void_p()
        _check(ctx.lib.bfj_buffer_alloc(ctx._h, self.nbytes, ctypes.byref(h)))
        self._ptr = h

    def data_ptr(self) -> int:
        return self._ptr.value or 0

    def stride(self, dim: int) -> int:
        return int(np.prod(self.shape[dim + 1:])) if dim < len(self.shape) else 1

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _check(self.ctx.lib.bfj_buffer_upload(self.ctx._h, self._ptr, a.ctypes.data, self.nbytes))
        return self

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx.lib.bfj_buffer_download(self.ctx._h, out.ctypes.data, self._ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "_ptr", None) and self._ptr.value and getattr(self.ctx, "_h", None):
            self.ctx.lib.bfj_buffer_free(self.ctx._h, self._ptr)
        self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def host_rows(data):
    """(array, n, d, ld) of a host float64 matrix as the C ABI takes it: rows of d values at
    a stride of ld >= d values, starting at the array's data pointer.  A view with unit
    column stride and a positive row stride (a column slice of a wider array, every k-th
    row) is passed as it is; anything else (reversed or broadcast rows, a column stride)
    is copied to a contiguous array first."""
    x = np.asarray(data)
    if x.dtype != np.float64 or x.ndim != 2:
        raise ValueError("data must be a 2-D float64 array")
    n, d = x.shape
    if n > 1 and (x.strides[1] != 8 or x.strides[0] < 8 * d or x.strides[0] % 8):
        x = np.ascontiguousarray(x)
    elif n <= 1 and x.strides[1] != 8 and d > 1:
        x = np.ascontiguousarray(x)
    ld = x.strides[0] // 8 if n > 1 else d
    return x, n, d, ld


def _dev_shape(x):
    """(n, d, ld, ptr) of a device matrix: a DeviceArray or any object with
    data_ptr()/shape/stride() (e.g. a torch tensor from the same HIP runtime)."""
    n, d = int(x.shape[0]), int(x.shape[1])
    if x.stride(1) != 1:
        raise ValueError("device rows need unit column stride")
    return n, d, (int(x.stride(0)) if n > 1 else d), int(x.data_ptr())


@dataclass
class JoinResult:
    counts: np.ndarray | None   # (n_eps, rows) int64, self excluded
    totals: np.ndarray          # (n_eps,) uint64, self included (reference totalNeighbors)
    stats: dict


class Context:
    """One HIP stream and the device buffers of the join (one per host thread)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib.bfj_init(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self.lib.bfj_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- single dataset -------------------------------------------------------------
    def self_join(self, data, eps, dtype="fp16", row_begin: int = 0, row_end: int | None = None,
                  counts: bool = True) -> JoinResult:
        """Join a host (numpy) float64 array of shape (n, d), or a strided view with unit column stride."""
        x, n, d, ld = host_rows(data)
        e = _eps_array(eps)
        if row_end is None:
            row_end = n
        rows = max(row_end - row_begin, 0)
        out = np.zeros((e.size, rows), dtype=np.int64) if counts else None
        tot = np.zeros(e.size, dtype=np.uint64)
        st = Stats()
        _check(self.lib.bfj_self_join(
            self._h, x.ctypes.data if n else None, n, d, ld, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            e.size, _dtype_id(dtype), row_begin, row_end, out.ctypes.data if counts and rows else None,
            tot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(st)))
        return JoinResult(out, tot, st.as_dict())

    def self_join_pairs(self, data, eps: float, dtype="fp16", capacity: int | None = None):
        """Result set of the join: every unordered pair (i, j), i < j, with
        sqrt(d2) <= eps, as a (m, 2) uint32 array in no particular order, plus the
        reference's ordered self-inclusive total and the pass statistics.  The buffer
        starts at ``capacity`` pairs (default 16 per row) and is grown once if the
        library reports that it needs more."""
        x = np.ascontiguousarray(data, dtype=np.float64)
        if x.ndim != 2:
            raise ValueError("data must be a 2-D float64 array")
        n, d = x.shape
        cap = int(capacity) if capacity is not None else max(16 * n, 1024)
        for _ in range(2):
            buf = np.zeros((max(cap, 1), 2), dtype=np.uint32)
            found, tot = ctypes.c_uint64(0), ctypes.c_uint64(0)
            st = Stats()
            rc = self.lib.bfj_self_join_pairs(self._h, x.ctypes.data if n else None, n, d, d, float(eps),
                                              _dtype_id(dtype), buf.ctypes.data, cap, ctypes.byref(found),
                                              ctypes.byref(tot), ctypes.byref(st))
            if rc == BFJ_ERR_CAPACITY and cap < found.value:
                cap = int(found.value)
                continue
            _check(rc)
            return buf[: found.value], int(tot.value), st.as_dict()
        raise JoinError(BFJ_ERR_CAPACITY, "pair buffer still too small after growing it")

    def reference_mode_join(self, data, eps: float, search_mode: int, compute_prec: int = 16,
                            accum_prec: int = 32) -> int:
        """The reference's own total for ``search_mode`` 1-4 (ordered pairs, self
        included), computed with its COMPUTE_PREC / ACCUM_PREC arithmetic (params.h:10-12)
        rather than the exact FP64 predicate; see ``bfj_reference_mode_join``."""
        x = np.ascontiguousarray(data, dtype=np.float64)
        if x.ndim != 2:
            raise ValueError("data must be a 2-D float64 array")
        n, d = x.shape
        tot = ctypes.c_uint64(0)
        _check(self.lib.bfj_reference_mode_join(self._h, x.ctypes.data if n else None, n, d, d, float(eps),
                                                int(search_mode), int(compute_prec), int(accum_prec),
                                                ctypes.byref(tot)))
        return int(tot.value)

    def to_device(self, host, dtype=np.float64) -> DeviceArray:
        a = np.ascontiguousarray(host, dtype=dtype)
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def self_join_device(self, x, eps, dtype="fp16", row_begin: int = 0, row_end: int | None = None,
                         counts=None) -> JoinResult:
        """Join device-resident float64 rows ``x`` (a :class:`DeviceArray` of shape (n, d), or
        another device matrix with unit column stride).

        ``counts``: optional int64 device buffer of shape (n_eps, rows) receiving the
        per-point counts (self excluded)."""
        if isinstance(x, DeviceArray) and x.dtype != np.float64:
            raise ValueError("rows must be float64")
        n, d, ld, ptr = _dev_shape(x)
        e = _eps_array(eps)
        if row_end is None:
            row_end = n
        if counts is not None:
            rows = row_end - row_begin
            if tuple(counts.shape) != (e.size, rows):
                raise ValueError("counts must have shape (n_eps, rows)")
        tot = np.zeros(e.size, dtype=np.uint64)
        st = Stats()
        _check(self.lib.bfj_self_join_device(
            self._h, ptr if n else None, n, d, ld, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            e.size, _dtype_id(dtype), row_begin, row_end, counts.data_ptr() if counts is not None else None,
            tot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(st)))
        return JoinResult(None, tot, st.as_dict())

    # -- batches ---------------------------------------------------------------------
    def _datasets(self, sets, device_resident: bool):
        arr = (Dataset * len(sets))()
        keep = []
        for k, s in enumerate(sets):
            data, eps = s["data"], _eps_array(s["eps"])
            tot = np.zeros(eps.size, dtype=np.uint64)
            if device_resident:
                n, d, ld, ptr = _dev_shape(data)
                cnt = s.get("counts")
                cptr = cnt.data_ptr() if cnt is not None else None
            else:
                data = np.ascontiguousarray(data, dtype=np.float64)
                n, d = data.shape
                ptr, ld = data.ctypes.data, d
                cnt = np.zeros((eps.size, n), dtype=np.int64)
                cptr = cnt.ctypes.data
            arr[k] = Dataset(ptr, n, d, ld, eps.size, eps.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cptr,
                             tot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
            keep.append((data, eps, tot, cnt))
        return arr, keep

    def self_join_batch(self, sets, dtype="fp16"):
        """sets: list of {"data": (n, d) float64 ndarray, "eps": [...]}.
        Returns (list of (counts, totals), stats)."""
        arr, keep = self._datasets(sets, device_resident=False)
        st = Stats()
        _check(self.lib.bfj_self_join_batch(self._h, arr, len(sets), _dtype_id(dtype), 0, ctypes.byref(st)))
        return [(c, t) for (_, _, t, c) in keep], st.as_dict()

    def plan(self, sets, dtype="fp16") -> "Plan":
        """sets: list of {"data": device float64 (n, d), "eps": [...], "counts": optional int64
        device buffer (n_eps, n)}."""
        return Plan(self, sets, dtype)

    # -- tiles -----------------------------------------------------------------------------
    def autotune(self, dtype, d: int, n: int = 32768) -> int:
        r = ctypes.c_int32(0)
        _check(self.lib.bfj_autotune(self._h, _dtype_id(dtype), d, n, ctypes.byref(r)))
        return r.value

    def set_tile_rows(self, dtype, rows: int):
        _check(self.lib.bfj_set_tile_rows(self._h, _dtype_id(dtype), rows))

    def set_capacity(self, candidate_chunks: int = 0, band_records: int = 0):
        """Capacities of the candidate stream (chunks of CHUNK_RECORDS records) and the band lists for
        the next pass (0 leaves one as it is).  A pass that needs more grows them and is
        repeated, so results never depend on these."""
        _check(self.lib.bfj_set_capacity(self._h, int(candidate_chunks), int(band_records)))

    def set_stream_limit(self, nbytes: int = 0):
        """Device memory the candidate stream and band lists of a pass may take (0: half of
        the device's memory).  A pass that would need more runs the join in slices, each
        followed by its band tests; the counts are the same (stats ``slices`` > 1)."""
        _check(self.lib.bfj_set_stream_limit(self._h, int(nbytes)))

    def pass_errors(self, max_candidates: int = 1 << 24, sample_tiles: int = 4096, seed: int = 12345) -> dict:
        """Accuracy of the last pass on this context (call it right after the pass): max
        |d2_approx - d2_fp64| over its candidate records (every k-th chunk beyond
        ``max_candidates``, 0 = all) and over ``sample_tiles`` seeded tiles of all pairs, the
        accumulation and quantisation bound ratios, and the pairs outside the certified
        interval (``violations``, 0 when the bound holds)."""
        e = Errors()
        _check(self.lib.bfj_pass_errors(self._h, int(max_candidates), int(sample_tiles), int(seed), ctypes.byref(e)))
        return e.as_dict()

    def set_reorder(self, enable: bool):
        """Turn the locality order of whole-dataset passes on (default) or off; results
        are the same either way."""
        _check(self.lib.bfj_set_reorder(self._h, 1 if enable else 0))


class Plan:
    """A batch captured once into a hipGraph; :meth:`run` replays it."""

    def __init__(self, ctx: Context, sets, dtype="fp16"):
        self.ctx = ctx
        self._arr, self._keep = ctx._datasets(sets, device_resident=True)
        h = ctypes.c_void_p()
        _check(ctx.lib.bfj_plan_create(ctx._h, self._arr, len(sets), _dtype_id(dtype), ctypes.byref(h)))
        self._h = h

    def run(self) -> dict:
        st = Stats()
        _check(self.ctx.lib.bfj_plan_run(self._h, ctypes.byref(st)))
        return st.as_dict()

    def totals(self):
        return [t.copy() for (_, _, t, _) in self._keep]

    def tiles(self) -> list[dict]:
        """The join shape of each K-chunk group (bfj_plan_tiles)."""
        arrs = [(ctypes.c_int32 * 33)() for _ in range(4)]
        n = ctypes.c_int32(0)
        _check(self.ctx.lib.bfj_plan_tiles(self._h, *arrs, 33, ctypes.byref(n)))
        keys = ("kc", "rows", "blocks_per_cu", "panel_bytes")
        return [{k: int(a[g]) for k, a in zip(keys, arrs)} for g in range(min(n.value, 33))]

    def close(self):
        if getattr(self, "_h", None):
            self.ctx.lib.bfj_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
