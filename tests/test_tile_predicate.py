"""The full-tile predicate of the tiled kernels near the 32-bit sample limit, and the XCD-aware unit order.

tests/cpp/tile_predicate.hip holds static_asserts over aeth::units::full_tile (the predicate
that picks a tile's unpredicated 16-byte path) at n = 2^32-4 and n = 2^32-1, where the last tile's
end passes 2^32; compiling it for gfx950 is the test (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_full_tile_predicate_does_not_wrap():
    src = os.path.join(ROOT, "tests", "cpp", "tile_predicate.hip")
    res = subprocess.run([HIPCC, "-fsyntax-only", "--offload-arch=gfx950", "-std=c++20", src],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
