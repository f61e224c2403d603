"""The reference's host API used as the reference's own caller uses it.

tests/dropin/dropin_driver is compiled against include/ with the reference's
configuration for its sample files (INPUT_DATA_DIM 18, COMPUTE_DIM 32, ACCUM_PREC 32):
it parses eps into ACCUM_TYPE (main.cu:20), reads the file with the reference-typed
importDataset (dataset.h:8-10), pads the rows as main.cu:39-57 does and calls
GPUJoinMainBruteForce (gpu_join.h:36-42) for each search mode.  The totals must equal
the exact counts of tests/golden/expected_counts.json at the binary32 eps ("f32"), and
the CLI, given the same eps string, must print the same totals (one rule for eps on
both surfaces: rounded to ACCUM_TYPE, as the reference's main.cu does).
"""
import json
import os
import subprocess

import pytest

import tensor_bf_main_amd as bfj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "dropin", "dropin_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")
MODES = [1, 2, 3, 4, 7, 8, 9, 10]

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="tests/dropin/dropin_driver not built")


def _gold(name):
    return json.load(open(os.path.join(GOLDEN, "expected_counts.json")))[name]


@needs_driver
def test_unknown_mode_exits_before_device_work():
    """gpu_join.cu:311-315: an unknown search mode prints an error and exits(1)."""
    r = subprocess.run([DRIVER, os.path.join(GOLDEN, "test100.txt"), "0.05", "42"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Unknown search mode" in r.stderr
    assert "rows 100 dim 18 compute_dim 32" in r.stdout  # the reader ran with the reference's dimension


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("name", ["test100.txt", "test128.txt"])
def test_reference_signature_totals(name):
    gold = _gold(name)["f32"]
    for eps_arg, total in zip(gold["eps_arg"], gold["totals"]):
        r = subprocess.run([DRIVER, os.path.join(GOLDEN, name), repr(eps_arg), *map(str, MODES)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = {int(ln.split()[1]): int(ln.split()[3]) for ln in r.stdout.splitlines() if ln.startswith("mode ")}
        assert got == {m: total for m in MODES}, (eps_arg, got)


@pytest.mark.gpu
@pytest.mark.parametrize("eps_arg", ["0.05", "0.1"])
def test_cli_takes_eps_as_the_reference_does(tmp_path, eps_arg):
    """bfjoin_main rounds eps to ACCUM_TYPE like main.cu:20, so it prints the drop-in
    surface's total; --eps-double keeps the double."""
    gold = _gold("test128.txt")
    want = gold["f32"]["totals"][gold["f32"]["eps_arg"].index(float(eps_arg))]
    for extra in ([], ["--eps-double"]):
        r = subprocess.run([bfj.CLI_PATH, os.path.join(GOLDEN, "test128.txt"), eps_arg, "4", "0", "--dim", "18",
                            "--log", str(tmp_path / "log.csv"), *extra], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        want_here = want if not extra else gold["totals"][gold["eps"].index(float(eps_arg))]
        assert f"Total result set size: {want_here}" in r.stdout
