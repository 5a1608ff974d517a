"""GPU: residency on one MI355X (include/aether/residency.hpp over the seam), with device lists that
repeat device 0 -- the single-GPU stand-in the reference's own tests use for several devices
(tests/test_PartitionedMove.cu, test_PeerAccess.cu: forced staging on one device).

* shards: the fused FP64 expression over a PartitionedArray of 2 and 4 ranks of 2^21 samples (whole
  groups of the sum tree), each rank's fused Neumaier sums combined over ranks in the tree's order, equal
  the one-array sums bit for bit, and gather() returns the one-array output bit for bit;
* a short last block of a 3-component array round-trips through scatter, a device update and an
  asynchronous gather;
* routes (SAME, STAGED when forced, HOST), staged copies (blocking, asynchronous, ordered with pageable
  memory) bit-identical to direct ones, moveTo / exchange through a staging transport, replica
  broadcasts, and the refusals.
Runs tests/cpp/residency_check.cpp."""
import os
import subprocess

import pytest

from aether_amd import build as B

pytestmark = pytest.mark.gpu


def run(*args):
    assert os.path.exists(B.RESIDENCY_CHECK), "residency_check not built (run the package build)"
    r = subprocess.run([B.RESIDENCY_CHECK, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


@pytest.mark.parametrize("parts", [2, 4])
def test_sharded_sums_equal_one_array(dev, parts):
    run("shards", parts)


def test_short_last_block_round_trip(dev):
    run("short")


def test_routes_transports_and_replicas(dev):
    run("transport")
