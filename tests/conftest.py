"""Shared test setup: the `gpu` marker, the oracle, and GPU helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def dev():
    """The GPU under test; fails (never skips) when the HIP library cannot run."""
    import torch

    import aether_amd
    from aether_amd import _lib

    _lib.lib()
    assert torch.cuda.is_available(), "gpu test without a GPU"
    torch.cuda.init()
    return torch.device("cuda:0")
