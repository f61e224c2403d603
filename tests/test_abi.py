"""C ABI and host-side checks that need no GPU.

* libbfjoin.so loads and exports every function include/bfjoin.h declares;
  the CLI exports the reference entry points of include/gpu_join.h and dataset.h.
* Argument validation and error reporting through the ABI.
* The C++ reader (the reference's dataset.cu rules) against the reference's own
  reader compiled into oracle/_ref, and the CLI's error behaviour (main.cu:62-108).
"""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import tensor_bf_main_amd as bfj
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bfjoin.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bfj_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = bfj.load_library()
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert lib.bfj_abi_version() == bfj.ABI_VERSION == 5


def test_cli_exports_reference_entry_points():
    out = subprocess.run(["nm", "-C", bfj.CLI_PATH], capture_output=True, text=True, check=True).stdout
    for sym in ("GPUJoinMainBruteForce(", "GPUJoinMainBruteForceNvidia(", "GPUJoinMainBruteForceEx(",
                "importDataset(", "importDatasetDim(", "importDatasetBinary(", "exportDatasetBinary("):
        assert sym in out, sym


def test_chunk_size_mirror_matches_the_header():
    txt = open(HEADER).read()
    m = re.search(r"^#define BFJ_CHUNK_RECORDS (\d+)", txt, re.M)
    assert m and int(m.group(1)) == bfj.CHUNK_RECORDS


def test_stats_mirror_matches_the_header():
    """The ctypes Stats layout is the C struct's: size and every field offset, as gcc
    lays out bfj_stats from include/bfjoin.h."""
    fields = [name for name, _ in bfj.Stats._fields_]
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"bfjoin.h\"\nint main(void) {\n"
    src += '    printf("%zu\\n", sizeof(bfj_stats));\n'
    src += "".join(f'    printf("%zu\\n", offsetof(bfj_stats, {f}));\n' for f in fields) + "    return 0;\n}\n"
    with tempfile.TemporaryDirectory() as td:
        c, exe = os.path.join(td, "layout.c"), os.path.join(td, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got[0] == ctypes.sizeof(bfj.Stats)
    assert got[1:] == [getattr(bfj.Stats, f).offset for f in fields]


def test_validation_without_device():
    lib = bfj.load_library()
    st = bfj.Stats()
    eps = (ctypes.c_double * 1)(0.5)
    # NULL context is rejected before anything touches a device.
    rc = lib.bfj_self_join(None, None, 10, 3, 3, eps, 1, 1, 0, 10, None, None, ctypes.byref(st))
    assert rc == 1 and b"ctx" in lib.bfj_last_error()
    rc = lib.bfj_self_join_batch(None, None, 0, 1, 0, None)
    assert rc == 1
    assert lib.bfj_set_stream_limit(None, 1 << 30) == 1 and b"ctx" in lib.bfj_last_error()
    assert lib.bfj_set_capacity(None, 1, 1) == 1
    n = ctypes.c_int(-1)
    assert lib.bfj_device_count(ctypes.byref(n)) == 0 and n.value >= 0
    if n.value == 0:
        h = ctypes.c_void_p()
        assert lib.bfj_init(0, ctypes.byref(h)) == 2
        assert b"device" in lib.bfj_last_error()
        with pytest.raises(bfj.JoinError) as ei:
            bfj.Context(0)
        assert ei.value.code == 2


def test_host_rows_keeps_views_and_copies_the_rest():
    """The host-buffer entry points take rows at a stride ld >= d from the array's data
    pointer: column slices and every-k-th-row views go through as they are (no copy);
    reversed, broadcast or column-strided arrays are copied first, never passed with a
    stride the library would read outside the buffer with."""
    wide = np.arange(70.0 * 50).reshape(70, 50)
    for view, ld in ((wide[:, 10:50], 50), (wide[::3, :17], 150), (wide, 50)):
        x, n, d, ld_got = bfj.host_rows(view)
        assert x is view and (n, d) == view.shape and ld_got == ld
        assert x.ctypes.data + ((n - 1) * ld_got + d) * 8 <= wide.ctypes.data + wide.nbytes
    for view in (wide[::-1], wide[:, ::2], np.broadcast_to(wide[0], (5, 50))):
        x, n, d, ld = bfj.host_rows(view)
        assert x.flags.c_contiguous and ld == d and np.array_equal(x, view)
    x, n, d, ld = bfj.host_rows(wide[3:4, 5:9])
    assert (n, d, ld) == (1, 4, 4)
    with pytest.raises(ValueError):
        bfj.host_rows(wide.astype(np.float32))


def test_tile_options_and_errors():
    assert bfj.tile_rows_options("fp16", 128) == [16, 32, 64, 96, 128]
    assert bfj.tile_rows_options("fp64", 128) == [16, 32]
    assert bfj.tile_rows_options("fp8", 256) == [16, 32, 64, 96, 128]
    assert bfj.tile_rows_options("bf16", 200) == [32, 64]
    for d in (1, 7, 31, 33, 64, 65, 96, 129, 200, 255, 256):
        for dt in bfj.DTYPES:
            assert bfj.tile_rows_options(dt, d), (dt, d)
    with pytest.raises(bfj.JoinError) as ei:
        bfj.tile_rows_options("fp16", 300)
    assert ei.value.code == 4
    with pytest.raises(bfj.JoinError) as ei:
        bfj.tile_rows_options(9, 16)
    assert ei.value.code == 1


def _cli(*args):
    return subprocess.run([bfj.CLI_PATH, *args], capture_output=True, text=True)


def test_cli_modes_5_and_6_are_errors():
    f = os.path.join(ROOT, "tests", "golden", "test100.txt")
    r = _cli(f, "0.05", "5", "--dim", "18")
    assert r.returncode == 1 and "removed" in r.stderr
    r = _cli(f, "0.05", "6", "--dim", "18")
    assert r.returncode == 1 and "Work in Progress" in r.stderr
    r = _cli(f, "0.05", "42", "--dim", "18")
    assert r.returncode == 1 and "Unknown search mode" in r.stderr


def _cli_rows(path, dim, *extra):
    r = _cli(path, "0.1", "4", "--dim", str(dim), "--parse-only", *extra)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    hdr = [ln for ln in lines if ln.startswith("[Main] ~ rows")][0].split()
    m, d = int(hdr[3]), int(hdr[5])
    body = lines[lines.index([ln for ln in lines if ln.startswith("[Main] ~ rows")][0]) + 1:]
    return np.array([[float(v) for v in ln.split()] for ln in body[:m]]).reshape(m, d)


def test_cli_reader_matches_oracle_restatement_and_reference():
    rng = np.random.default_rng(12)
    vals = rng.random(90 * 4 + 11)
    cases = [", ".join(f"{v:.6f}" for v in vals),
             "\n".join(" ".join(f"{v:.9g}" for v in vals[k:k + 13]) for k in range(0, vals.size, 13)),
             ",".join(f"{v:.4e}" for v in vals) + "\nnot-a-number 1 2 3"]
    for txt in cases:
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
            f.write(txt)
            path = f.name
        try:
            mine = _cli_rows(path, 90)
            np.testing.assert_array_equal(mine, oracle.parse_text(txt, 90))
            ref = oracle.ref_import(path)
            if ref is not None:
                np.testing.assert_array_equal(mine, ref)
        finally:
            os.unlink(path)
    gold = _cli_rows(os.path.join(ROOT, "tests", "golden", "test128.txt"), 18)
    assert gold.shape == (128, 18)


def test_cli_binary_format_roundtrip():
    rng = np.random.default_rng(13)
    x = rng.standard_normal((37, 5))
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(b"BFJDATA1" + np.array([37, 5], dtype="<i8").tobytes() + x.astype("<f8").tobytes())
        path = f.name
    try:
        got = _cli_rows(path, 18, "--binary")
        np.testing.assert_array_equal(got, x)
    finally:
        os.unlink(path)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and exits before any GPU work."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_profile_traffic_lines_parse():
    """bench.py's roofline.traffic comes from the profiles' machine-readable lines: each one
    parses to a positive byte count, and the headline configuration (workload 3, FP16,
    R = 128) and the FP64 workload 2 resolve to theirs."""
    import importlib.util
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    prof = os.path.join(root, "profiles")
    lines = [ln.strip() for name in sorted(os.listdir(prof)) if name.endswith(".md")
             for ln in open(os.path.join(prof, name)) if ln.startswith("machine-readable:")]
    assert lines
    pat = re.compile(r"^machine-readable: workload=(\d) dtype=(\w+) tile_rows=(\d+) traffic_bytes=(\d+)$")
    for ln in lines:
        m = pat.match(ln)
        assert m, ln
        w, dt, r, b = int(m.group(1)), m.group(2), int(m.group(3)), int(m.group(4))
        assert b > 0
        assert bench.measured_traffic(w, dt, r) == float(b)
    assert bench.measured_traffic(3, "fp16", 128) > 1e9
    assert bench.measured_traffic(2, "fp64", 16) > 1e8
    assert bench.measured_traffic(3, "fp16", 64) is None
