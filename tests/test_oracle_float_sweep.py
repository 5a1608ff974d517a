"""Pins the oracle's float math against long-double libm with oracle/float_sweep.c (built by
`make -C oracle float_sweep`): every 1009th float argument here (the stride-1 run visits them all), and
seeded pow pairs. Bounds: what the stride-1 run measured, rounded up (sin/cos 0.667, exp 0.780 with
subnormal results, log 0.643, tan 0.783, atan/asin/acos 0.501, tanh 0.640, pow 0.792 with subnormal
results)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOUND = {"sin": 0.67, "cos": 0.67, "exp": 0.78, "log": 0.65, "tan": 0.79, "atan": 0.502, "asin": 0.502,
         "acos": 0.502, "tanh": 0.65, "pow0": 0.8, "pow1": 0.8, "pow2": 0.8}


@pytest.fixture(scope="module")
def sweep(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sweep") / "float_sweep")  # one build per test process
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"SWEEP={exe}", exe], check=True)
    out = subprocess.run([exe, "1009"], check=True, capture_output=True, text=True, timeout=300).stdout
    rows = {}
    for line in out.splitlines():
        m = re.match(r"(\w+)\s+max ([0-9.]+) ulp .* \(([0-9.]+)\)$", line)
        if m:
            rows[m.group(1)] = (float(m.group(2)), float(m.group(3)))
    return rows


@pytest.mark.parametrize("fn", sorted(BOUND))
def test_float_routine_within_stated_bound(sweep, fn):
    worst, not_cr = sweep[fn]
    assert worst <= BOUND[fn], (fn, worst)
    assert not_cr <= 0.02, (fn, not_cr)
