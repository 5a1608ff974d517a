"""GPU: the on-disk cache of runtime-compiled tape kernels (AETHER_KERNEL_CACHE).

A first process compiles the fused benchmark tree (with sums) and a Vec3 tree and leaves their code objects in
the cache directory; a second process with the same directory compiles nothing, loads them, and writes the
same values and sums; a corrupted file is compiled again and replaced. Each process runs in its own
interpreter because the in-process cache would otherwise answer first."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import ctypes as C, hashlib, sys
sys.path.insert(0, ROOT)
import torch
import aether_amd as ae
from aether_amd import expr as E
n = 100_003
arrs = []
for s in (11, 12, 13):
    a = ae.Array(n, 3, torch.float64); ae.array_fill_uniform(a, s, -1.0, 1.0); arrs.append(a)
a, b, c = arrs
d, e = ae.Array(n, 3, torch.float64), ae.Array(n, 1, torch.float64)
sums = E.assign(d, a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)), ae.SUM_NEUMAIER)
E.assign(e, a.cross(b).dot(c) + E.sqrt(a.dot(a) + 1.0))
torch.cuda.synchronize()
h = hashlib.sha256()
h.update(d.data[:, :n].cpu().numpy().tobytes()); h.update(e.data[:, :n].cpu().numpy().tobytes())
h.update(sums.cpu().numpy().tobytes())
cc, hh, ff = C.c_uint64(), C.c_uint64(), C.c_uint64()
ae._lib.call("aeth_tape_jit_stats", C.byref(cc), C.byref(hh), C.byref(ff))
print(cc.value, hh.value, ff.value, h.hexdigest())
""".replace("ROOT", repr(ROOT))


def run(cache_dir):
    env = dict(os.environ, AETHER_KERNEL_CACHE=str(cache_dir))
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, env=env, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    compiled, hits, fallbacks, digest = r.stdout.split()[-4:]
    return int(compiled), int(hits), int(fallbacks), digest


def test_second_process_loads_instead_of_compiling(dev, tmp_path):
    c1, h1, f1, d1 = run(tmp_path)
    files = sorted(p for p in os.listdir(tmp_path) if p.endswith(".co"))
    assert c1 >= 2 and f1 == 0 and len(files) >= 2, (c1, f1, files)
    assert not [p for p in os.listdir(tmp_path) if ".tmp" in p]
    c2, h2, f2, d2 = run(tmp_path)
    assert c2 == 0 and f2 == 0 and h2 >= 2, (c2, h2, f2)
    assert d2 == d1
    # a damaged code object is compiled again and replaced
    victim = tmp_path / files[0]
    victim.write_bytes(b"not a code object")
    c3, _, f3, d3 = run(tmp_path)
    assert c3 == 1 and f3 == 0 and d3 == d1, (c3, f3)
    assert victim.stat().st_size > 1000
