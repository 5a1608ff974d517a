"""GPU: generic tape lowering (aeth_eval_tape) agrees bit for bit with the
specialised kernels for the same trees, with the oracle for other trees, and
with the array fills for distribution leaves.
"""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def filled(n, dt, seeds=(42, 43, 44), ncomp=3):
    out = []
    for s in seeds:
        a = ae.Array(n, ncomp, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        out.append(a)
    return out


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [1, 2047, 2048, 70001, 1 << 22])
@pytest.mark.parametrize("mode", [ae.SUM_NONE, ae.SUM_NEUMAIER, ae.SUM_PLAIN])
def test_cfg0_tape_equals_specialised(dev, dt, n, mode):
    a, b, c = filled(n, dt)
    d1, d2 = ae.Array(n, 3, dt), ae.Array(n, 3, dt)
    s1 = ae.eval_cfg0(a, b, c, d1, mode)
    s2 = E.assign(d2, a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)), mode)
    torch.cuda.synchronize()
    assert torch.equal(d1.data[:, :n], d2.data[:, :n])
    if mode != ae.SUM_NONE:
        assert torch.equal(s1, s2)


def test_vec3_tape_equals_specialised(dev):
    n = 100_003
    a, b, c = filled(n, torch.float64)
    e1 = torch.empty(n, dtype=torch.float64, device="cuda")
    ae.eval_vec3geo(a, b, c, e1)
    e2 = ae.Array(n, 1, torch.float64)
    E.assign(e2, a.cross(b).dot(c) + E.sqrt(a.dot(a) + 1.0))
    torch.cuda.synchronize()
    assert torch.equal(e1, e2.data[0, :n])


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_random_leaves_equal_fills(dev, dt):
    n = 50_000
    g = ae.Generator.from_seed(7)
    x1, x2 = ae.Array(n, 3, dt), ae.Array(n, 3, dt)
    ae.array_fill_uniform(x1, g.uniform_seed(), -2.0, 5.0)
    E.assign(x2, E.uniform(g, 3, -2.0, 5.0))
    y1, y2 = ae.Array(n, 3, dt), ae.Array(n, 3, dt)
    ae.array_fill_normal(y1, g.normal_seed(), 1.0, 0.5)
    E.assign(y2, E.normal(g, 3, 1.0, 0.5))
    torch.cuda.synchronize()
    assert torch.equal(x1.data[:, :n], x2.data[:, :n])
    assert torch.equal(y1.data[:, :n], y2.data[:, :n])


def test_other_ops_vs_oracle_math(dev, orc):
    n = 40_000
    (a,) = filled(n, torch.float64, seeds=(5,), ncomp=1)
    d = ae.Array(n, 1, torch.float64)
    E.assign(d, E.log(E.abs_(a) + 0.5) / 3.0 - E.cos(a) * E.sqrt(a * a + 1.0) / (a + 2.0))
    torch.cuda.synchronize()
    x = a.data[0, :n].cpu().numpy()
    third = 1.0 / 3.0
    ref = orc.math64("log", np.abs(x) + 0.5) * third - orc.math64("cos", x) * np.sqrt(x * x + 1.0) / (x + 2.0)
    assert np.array_equal(d.data[0, :n].cpu().numpy(), ref)


def test_in_place_accumulate(dev):
    n = 10_000
    (a,) = filled(n, torch.float64, seeds=(3,))
    before = a.data[:, :n].clone()
    E.assign(a, a * 2.0 + 1.0)
    torch.cuda.synchronize()
    assert torch.equal(a.data[:, :n], before * 2.0 + 1.0)


def test_invalid_tape_is_rejected(dev):
    n = 100
    (a,) = filled(n, torch.float64, seeds=(3,), ncomp=1)
    d = ae.Array(n, 1, torch.float64)
    progs, leaves, consts, rngs = E.lower(d, a + 1.0)
    progs[0].n_insns = 2  # drops the ADD: two values left on the stack
    import ctypes as C

    ptrs = (C.c_void_p * len(leaves))(*[lf.ptr() for lf in leaves])
    pitch = (C.c_uint64 * len(leaves))(*[lf.capacity for lf in leaves])
    cst = (C.c_double * 1)(*consts)
    with pytest.raises(ae.AetherError):
        ae._lib.call("aeth_eval_tape", ae.F64, 1, progs, len(leaves), ptrs, pitch, 1, cst, 0, None, n, -1, None,
                     ae.stream_handle())


@pytest.mark.parametrize("n", [1000, 70001])
def test_cross_component_aliasing_is_sequential(dev, n):
    # a = a x b: aether stores component 0 of a sample before it evaluates component 1
    # (expr/Assign.h:124-160), so components 1 and 2 read the updated components
    a, b = filled(n, torch.float64, seeds=(3, 4))
    A = a.data[:, :n].cpu().numpy().copy()
    B = b.data[:, :n].cpu().numpy()
    sums = E.assign(a, a.cross(b), ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    A[0] = A[1] * B[2] - A[2] * B[1]
    A[1] = A[2] * B[0] - A[0] * B[2]
    A[2] = A[0] * B[1] - A[1] * B[0]
    assert np.array_equal(a.data[:, :n].cpu().numpy(), A)
    assert torch.equal(sums, ae.reduce(a, ae.SUM_NEUMAIER))


def test_two_programs_one_plane_rejected(dev):
    n = 100
    a, b = filled(n, torch.float64, seeds=(3, 4))
    progs, leaves, consts, rngs = E.lower(b, a * 2.0)
    progs[1].out_comp = 0  # programs 0 and 1 both write plane 0 of b
    import ctypes as C

    ptrs = (C.c_void_p * len(leaves))(*[lf.ptr() for lf in leaves])
    pitch = (C.c_uint64 * len(leaves))(*[lf.capacity for lf in leaves])
    cst = (C.c_double * len(consts))(*consts)
    with pytest.raises(ae.AetherError):
        ae._lib.call("aeth_eval_tape", ae.F64, 3, progs, len(leaves), ptrs, pitch, len(consts), cst, 0, None, n, -1,
                     None, ae.stream_handle())


def test_runtime_lowering_was_used(dev):
    import ctypes as C

    c, h, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
    ae._lib.call("aeth_tape_jit_stats", C.byref(c), C.byref(h), C.byref(f))
    assert c.value >= 1 and f.value == 0


INTERP_SCRIPT = r"""
import torch, aether_amd as ae
from aether_amd import expr as E
n = 30001
arrs = []
for s in (42, 43, 44):
    a = ae.Array(n, 3, torch.float64); ae.array_fill_uniform(a, s); arrs.append(a)
a, b, c = arrs
d1, d2 = ae.Array(n, 3, torch.float64), ae.Array(n, 3, torch.float64)
s1 = ae.eval_cfg0(a, b, c, d1, ae.SUM_NEUMAIER)
s2 = E.assign(d2, a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)), ae.SUM_NEUMAIER)
torch.cuda.synchronize()
assert torch.equal(d1.data[:, :n], d2.data[:, :n]) and torch.equal(s1, s2)
A = a.data[:, :n].cpu().numpy().copy(); B = b.data[:, :n].cpu().numpy()
E.assign(a, a.cross(b)); torch.cuda.synchronize()
A[0] = A[1] * B[2] - A[2] * B[1]; A[1] = A[2] * B[0] - A[0] * B[2]; A[2] = A[0] * B[1] - A[1] * B[0]
import numpy as np
assert np.array_equal(a.data[:, :n].cpu().numpy(), A)
print("interpreter ok")
"""


def test_interpreter_path_same_bits(dev):
    import os
    import subprocess
    import sys

    env = dict(os.environ, AETHER_TAPE_INTERPRET="1")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", INTERP_SCRIPT], cwd=root, env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and "interpreter ok" in r.stdout, r.stderr[-2000:]
