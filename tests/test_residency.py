"""Residency on the host (no GPU): the partition algebra of include/aether/residency.hpp
(padded_block_samples, partition_view: residency/Partition.h:52-160) against the reference's own rows
(tests/test_Partition.cpp: 100 samples in 4 blocks of pitch 32 with a short last block, 96 in 3 exact
blocks, the refusals), host routes (PeerAccess.h: HOST unless both places are GPU memory), an empty
device list refused, and the rank combine's halving order. Runs tests/cpp/residency_check.cpp."""
import os
import subprocess

from aether_amd import build as B


def test_partition_algebra_and_host_routes():
    assert os.path.exists(B.RESIDENCY_CHECK), "residency_check not built (run the package build)"
    r = subprocess.run([B.RESIDENCY_CHECK, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
