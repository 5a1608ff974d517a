"""GPU: the sticky device error word (err/DeviceFlag.h:86-160) with the reference's own cases
(tests/test_DeviceFlag.cu): initially clear; set from a kernel and not cleared by a read; checkAndClear
throws once and clears; clear; 64 x 256 lanes racing to set it leave exactly 1; set only by the lane that
meets a bad input and sticky across a later clean launch; moves; AppendCounter's overflow as the consumer.
The C++ cases run tests/cpp/flag_check.hip (its own kernels take DeviceFlagView by value); the Python
mirror's cases run here."""
import os
import subprocess

import pytest
import torch

import aether_amd as ae
from aether_amd import accum as A
from aether_amd import build as B

pytestmark = pytest.mark.gpu


def test_cpp_flag_cases(dev):
    assert os.path.exists(B.FLAG_CHECK), "flag_check not built (run the package build)"
    r = subprocess.run([B.FLAG_CHECK, "device"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_python_flag_set_by_an_overflowing_claim(dev):
    flag = A.DeviceFlag()
    assert not flag.is_set()
    flag.check_and_clear("clean")  # no raise
    # a claim past capacity sets the counter's flag from the device
    ctr = A.AppendCounter(10)
    x = torch.ones(1000, dtype=torch.float64, device="cuda")
    out = torch.empty(10, dtype=torch.int32, device="cuda")
    A.append_where(x, 0.5, ctr, out, "lane")
    assert ctr.overflow.is_set() and ctr.overflow.is_set()  # a read does not clear it
    with pytest.raises(ae.AetherError, match="step 3 failed on DeviceFlag"):
        ctr.check_overflow("step 3")
    assert not ctr.overflowed()
    ctr.check_overflow("step 4")  # cleared: no raise
    # the same flag passed to the C entry directly, then cleared without raising
    A.append_where(x, 0.5, ctr, out, "tile")
    assert ctr.overflowed()
    ctr.overflow.clear()
    assert not ctr.overflowed()
