"""GPU: the runtime evaluator over strided views (eval/Runtime.h) against numpy on the same inputs:
every op, both widths and the casts, dense and strided (transposed, sliced, stride-0 broadcast)
operands, ranks 0..6, DLPack producers, and the reference's validation errors (shape and dtype
mismatches, a cast outside ASSIGN). Nothing is contracted, so results are bit-exact with numpy."""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import runtime as R

pytestmark = pytest.mark.gpu

REF = {"assign": lambda a, b, s: a, "add": lambda a, b, s: a + b, "sub": lambda a, b, s: a - b,
       "scale": lambda a, b, s: s * a, "add_scaled": lambda a, b, s: a + s * b,
       "sub_scaled": lambda a, b, s: a - s * b}


def rnd(shape, dt, seed):
    return torch.from_numpy(np.random.default_rng(seed).uniform(-2, 2, shape)).to(dt)


@pytest.mark.parametrize("op", sorted(REF))
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(), (7,), (1000003,), (33, 65), (3, 4, 5, 6, 2, 3)])
def test_dense_ops_match_numpy(dev, op, dt, shape):
    a, b = rnd(shape, dt, 1), rnd(shape, dt, 2)
    s = 0.3
    got = R.runtime_eval(op, a.cuda(), b.cuda(), scalar=s)
    np_dt = np.float64 if dt == torch.float64 else np.float32
    want = REF[op](a.numpy(), b.numpy(), np_dt(s))
    assert np.array_equal(got.cpu().numpy(), np.asarray(want, np_dt))


@pytest.mark.parametrize("op", ["add_scaled", "scale"])
@pytest.mark.parametrize("dt,n", [(torch.float64, (1 << 25) + 3), (torch.float32, (1 << 26) + 1029),
                                  (torch.float64, 4 * 1024 * 2 + 2 * 256 * 2 + 1)])
def test_dense_units_and_streaming(dev, op, dt, n):
    """Sizes past the streaming threshold (three operands of 256 MiB) and around the dense kernel's 16 KiB
    units (whole units, spare vectors, a scalar tail), bit-exact against numpy."""
    a, b = rnd((n,), dt, 3), rnd((n,), dt, 4)
    s = -0.7
    got = R.runtime_eval(op, a.cuda(), b.cuda(), scalar=s).cpu().numpy()
    np_dt = np.float64 if dt == torch.float64 else np.float32
    want = np.asarray(REF[op](a.numpy(), b.numpy(), np_dt(s)), np_dt)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [5, 4096 + 3, (1 << 20) + 7])
def test_dense_in_place(dev, n):
    """out may be an input: a = a + s*b and b = s*b in place give the out-of-place bits."""
    a, b = rnd((n,), torch.float64, 5), rnd((n,), torch.float64, 6)
    want_a = np.asarray(a.numpy() + 0.25 * b.numpy())
    want_b = np.asarray(0.25 * b.numpy())
    ga, gb = a.cuda(), b.cuda()
    R.runtime_eval("add_scaled", ga, gb, out=ga, scalar=0.25)
    R.runtime_eval("scale", gb, out=gb, scalar=0.25)
    assert np.array_equal(ga.cpu().numpy(), want_a) and np.array_equal(gb.cpu().numpy(), want_b)


@pytest.mark.parametrize("op", ["add", "sub_scaled", "scale", "assign"])
def test_strided_operands(dev, op):
    base = rnd((64, 96), torch.float64, 3).cuda()
    a = base.t()                       # transposed
    b = rnd((96, 200), torch.float64, 4).cuda()[:, ::3][:, :64]   # sliced, stride 3
    out = torch.zeros((64, 96), dtype=torch.float64, device="cuda").t()  # written through a transpose
    R.runtime_eval(op, a, b, out, scalar=-1.5)
    want = REF[op](a.cpu().numpy(), b.cpu().numpy(), -1.5)
    assert np.array_equal(out.cpu().numpy(), want)


def test_broadcast_by_zero_stride(dev):
    row = rnd((1, 50), torch.float32, 5).cuda().expand(40, 50)  # stride 0 along rows
    a = rnd((40, 50), torch.float32, 6).cuda()
    got = R.runtime_eval("add", a, row)
    assert np.array_equal(got.cpu().numpy(), a.cpu().numpy() + row.cpu().numpy())


@pytest.mark.parametrize("src,dst", [(torch.float64, torch.float32), (torch.float32, torch.float64)])
def test_assign_casts(dev, src, dst):
    a = rnd((12345,), src, 7).cuda()
    out = torch.empty(12345, dtype=dst, device="cuda")
    R.runtime_eval("assign", a, out=out)
    assert np.array_equal(out.cpu().numpy(), a.cpu().numpy().astype(out.cpu().numpy().dtype))


def test_dlpack_producer(dev):
    class Producer:
        def __init__(self, t):
            self.t = t

        def __dlpack__(self, stream=None):
            return self.t.__dlpack__()

        def __dlpack_device__(self):
            return self.t.__dlpack_device__()

    a = rnd((100,), torch.float64, 8).cuda()
    got = R.runtime_eval("scale", Producer(a), scalar=2.0)
    assert torch.equal(got, a * 2.0)


def test_validation_errors(dev):
    a = torch.zeros(10, dtype=torch.float64, device="cuda")
    with pytest.raises(ae.AetherError, match="shape mismatch: a vs out"):
        R.runtime_eval("assign", a, out=torch.zeros(11, dtype=torch.float64, device="cuda"))
    with pytest.raises(ae.AetherError, match="shape mismatch: b vs out"):
        R.runtime_eval("add", a, torch.zeros(9, dtype=torch.float64, device="cuda"))
    with pytest.raises(ae.AetherError, match="only ASSIGN casts"):
        R.runtime_eval("scale", a, out=torch.zeros(10, dtype=torch.float32, device="cuda"))
    with pytest.raises(ae.AetherError, match="dtype mismatch: b vs out"):
        R.runtime_eval("add", a, torch.zeros(10, dtype=torch.float32, device="cuda"))
    with pytest.raises(ae.AetherError, match="double/float only"):
        R.view(torch.zeros(3, dtype=torch.int32, device="cuda"))
