"""The CPU oracle and the host driver / readers under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

* oracle/libbfj_oracle_san.so (bfj_oracle.c built with -fsanitize=address,undefined)
  runs the oracle's own test suites in a child process with the ASan runtime
  preloaded; any invalid access or undefined operation aborts the child.
* tensor_bf_main_amd/bfjoin_main_san (csrc/host/*.cpp with the same flags) reads the
  reference's sample files, text in the reference's formats and the binary format, and
  takes the error paths of main.cu:62-108, without a report.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SAN = os.path.join(ROOT, "oracle", "libbfj_oracle_san.so")
CLI_SAN = os.path.join(ROOT, "tensor_bf_main_amd", "bfjoin_main_san")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def _env(**extra):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env.update(extra)
    return env


def _clean(r):
    return "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr


@pytest.mark.skipif(not os.path.exists(ORACLE_SAN) or not _asan_runtime(), reason="sanitizer oracle build missing")
def test_oracle_suites_under_sanitizers():
    env = _env(LD_PRELOAD=_asan_runtime(), BFJ_ORACLE_LIB=ORACLE_SAN)
    # The runtime really is in the child, and the oracle loaded there is the instrumented one.
    probe = ("import ctypes, sys; sys.path.insert(0, %r); from oracle import oracle; "
             "assert hasattr(ctypes.CDLL(None), '__asan_init'); oracle._load(); "
             "assert oracle._LIB_PATH == %r; print('ok')") % (ROOT, ORACLE_SAN)
    r = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-3000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py"),
                        os.path.join(ROOT, "tests", "test_reference_modes.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and _clean(r), (r.stdout[-2000:], r.stderr[-3000:])


@pytest.mark.skipif(not os.path.exists(ORACLE_SAN) or not _asan_runtime(), reason="sanitizer oracle build missing")
def test_sanitizer_catches_a_planted_overread():
    """The harness can fail: a row buffer one value shorter than d, passed to the
    instrumented oracle in a child process, is reported as a heap overflow."""
    code = ("import ctypes, sys; sys.path.insert(0, %r); from oracle import oracle; lib = oracle._load(); "
            "m = ctypes.CDLL(None).malloc; m.restype = ctypes.c_void_p; m.argtypes = [ctypes.c_size_t]; "
            "q = (ctypes.c_double * 8)(*range(8)); "
            "c = ctypes.cast(m(7 * 8), ctypes.POINTER(ctypes.c_double)); "
            "lib.bfo_pair_d2(q, c, 8)") % ROOT
    env = _env(LD_PRELOAD=_asan_runtime(), BFJ_ORACLE_LIB=ORACLE_SAN)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]


def _cli(*args):
    return subprocess.run([CLI_SAN, *args], env=_env(), capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(not os.path.exists(CLI_SAN), reason="sanitizer CLI build missing")
def test_host_driver_and_readers_under_sanitizers():
    for name, n in (("test100.txt", 100), ("test128.txt", 128)):
        r = _cli(os.path.join(GOLDEN, name), "0.05", "4", "--dim", "18", "--parse-only")
        assert r.returncode == 0 and _clean(r) and f"rows {n} dim 18" in r.stdout, r.stderr[-2000:]
    rng = np.random.default_rng(3)
    vals = rng.random(90 * 3 + 7)
    texts = [", ".join(f"{v:.6f}" for v in vals), "\n".join(f"{v:.3e}" for v in vals) + "\n1 2 x 3", "", "abc"]
    with tempfile.TemporaryDirectory() as td:
        for k, txt in enumerate(texts):
            path = os.path.join(td, f"t{k}.txt")
            open(path, "w").write(txt)
            r = _cli(path, "0.1", "4", "--dim", "90", "--parse-only")
            assert r.returncode == 0 and _clean(r), (k, r.stderr[-2000:])
        x = rng.standard_normal((21, 7))
        path = os.path.join(td, "x.bin")
        open(path, "wb").write(b"BFJDATA1" + np.array([21, 7], dtype="<i8").tobytes() + x.astype("<f8").tobytes())
        r = _cli(path, "0.1", "4", "--binary", "--parse-only")
        assert r.returncode == 0 and _clean(r) and "rows 21 dim 7" in r.stdout
        bad = os.path.join(td, "bad.bin")
        open(bad, "wb").write(b"BFJDATA1" + np.array([1000, 7], dtype="<i8").tobytes() + b"\0" * 40)  # truncated
        r = _cli(bad, "0.1", "4", "--binary", "--parse-only")
        assert r.returncode == 1 and _clean(r)
        r = _cli(os.path.join(td, "missing.txt"), "0.1", "4", "--parse-only")
        assert r.returncode == 1 and _clean(r)
    for mode, msg in (("5", "removed"), ("6", "Work in Progress"), ("42", "Unknown search mode")):
        r = _cli(os.path.join(GOLDEN, "test100.txt"), "0.05", mode, "--dim", "18")
        assert r.returncode == 1 and msg in r.stderr and _clean(r)
