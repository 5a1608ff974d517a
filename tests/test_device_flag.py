"""DeviceFlagView's host arm (err/DeviceFlag.h:49-84): a relaxed atomic store of 1 into a word of host
memory, idempotent under 8 threads racing to set it. Runs tests/cpp/flag_check.hip's host mode (no GPU)."""
import os
import subprocess

from aether_amd import build as B


def test_host_arm_sets_and_races():
    assert os.path.exists(B.FLAG_CHECK), "flag_check not built (run the package build)"
    r = subprocess.run([B.FLAG_CHECK, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
