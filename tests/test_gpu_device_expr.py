"""GPU: the C++ front end's expressions evaluated per sample inside a caller's own HIP kernels.

tests/cpp/expr_device_check.hip launches kernels that write `dv[SampleIndex::make(threadIdx.x, blockIdx.x,
blockDim.x)] = expr` over device views (the reference's device spelling, expr/Assign.h:251-346 and
index/SampleIndex.h:46-50) and compares them bit for bit with the same expression assigned to the whole
view (`dv = expr`, the runtime-compiled kernel whose own parity with the oracle the tape tests hold): the
fused benchmark tree and the Vec3 tree in double and float, the elementwise functions, folds and norms,
quaternions, Philox distribution leaves, `+=` / `-=` and `get()`, at 1, 257 and 100003 samples."""
import os
import subprocess

import pytest

from aether_amd import build as B

pytestmark = pytest.mark.gpu


def test_device_expressions_equal_whole_array_assignment(dev):
    assert os.path.exists(B.EXPR_DEVICE_CHECK), "expr_device_check not built (run the package build)"
    r = subprocess.run([B.EXPR_DEVICE_CHECK], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
