"""GPU: aether_amd.dsc.fft -- the reference's aether.fft rows (dsc/tests/test_fft.py:78-268) over aether
arrays, on both the host ("cpu", staged through the device) and the device. numpy.fft is the oracle and the
bound is the reference's ULP-scaled one: the largest absolute error of a transform over N points is at most
(8 + 4 log2 N) * eps * ||x||_2 * scale. Plans, their errors, a plan captured into a graph and replayed, the
refusal of plan creation inside a capture, and bitwise determinism of repeated transforms."""
import math
import pathlib
import subprocess
import sys
import textwrap

import numpy as np
import pytest
import torch

import aether_amd.dsc as aether
from aether_amd.dsc import fft as F

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
DTYPES = ["float32", "float64"]
DEVICES = ["cpu", "cuda"]


def pairs(c, dtype):
    return np.ascontiguousarray(np.stack([c.real, c.imag], -1).astype(dtype))


def cplx(p):
    p = np.asarray(p, dtype=np.float64)
    return p[..., 0] + 1j * p[..., 1]


def tol(dtype, points, x, scale):
    return (8 + 4 * math.log2(points + 1)) * np.finfo(dtype).eps * float(np.linalg.norm(np.asarray(x, np.float64))) * scale


def rand_c(shape, dtype, seed=0):
    r = np.random.default_rng(seed)
    return pairs(r.uniform(-1, 1, shape) + 1j * r.uniform(-1, 1, shape), dtype)


def scale_of(norm, forward, n):
    if norm == "ortho":
        return 1 / math.sqrt(n)
    return 1 / n if forward == (norm == "forward") else 1.0


def to_host(a):
    return np.array(aether.asarray(a, device="cpu"), copy=True)


def on(x, device):
    return aether.asarray(x, device=device)


@pytest.mark.parametrize("device", DEVICES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_c2c_lengths_match_numpy(dev, device, dtype):
    for n in (1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 17, 31, 64, 100, 127, 1000):
        x = rand_c((3, n), dtype, n)
        for forward in (True, False):
            fn = F.fft if forward else F.ifft
            y = fn(on(x, device))
            assert y.device.kind == device
            got = to_host(y)
            ref = (np.fft.fft if forward else np.fft.ifft)(cplx(x), axis=-1)
            err = np.abs(cplx(got) - ref).max()
            assert err <= tol(dtype, n, x, scale_of("backward", forward, n)), (n, forward, err)


@pytest.mark.parametrize("device", DEVICES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_multi_axis_and_batched_match_numpy(dev, device, dtype):
    x = rand_c((4, 7, 6), dtype, 5)
    xc = cplx(x)
    cases = [((1,), np.fft.fft(xc, axis=1)), ((0, 2), np.fft.fftn(xc, axes=(0, 2))), ((0, 1, 2), np.fft.fftn(xc)),
             ((2, 0), np.fft.fftn(xc, axes=(2, 0)))]
    for axes, ref in cases:
        got = to_host(F.fftn(on(x, device), axes=axes))
        pts = int(np.prod([x.shape[a] for a in axes]))
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, pts, x, 1.0), axes
        back = to_host(F.ifftn(on(got, device), axes=axes))
        assert np.abs(cplx(back) - xc).max() <= 2 * tol(dtype, pts, got, 1.0 / pts), axes


@pytest.mark.parametrize("device", DEVICES)
@pytest.mark.parametrize("norm", ["backward", "forward", "ortho"])
def test_norms_match_numpy(dev, device, norm):
    x = rand_c((2, 12), "float64", 9)
    for forward in (True, False):
        fn = F.fft if forward else F.ifft
        got = to_host(fn(on(x, device), norm=norm))
        ref = (np.fft.fft if forward else np.fft.ifft)(cplx(x), norm=norm)
        assert np.abs(cplx(got) - ref).max() <= tol("float64", 12, x, scale_of(norm, forward, 12))


@pytest.mark.parametrize("device", DEVICES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_rfft_irfft_match_numpy(dev, device, dtype):
    for n in (2, 3, 7, 8, 9, 16, 17, 50, 128):
        r = np.random.default_rng(n).uniform(-1, 1, (3, n)).astype(dtype)
        got = to_host(F.rfft(on(r, device)))
        ref = np.fft.rfft(r.astype(np.float64), axis=-1)
        assert got.shape == ref.shape + (2,)
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, n, r, 1.0), n
        back = to_host(F.irfft(on(got, device), n=n))
        assert np.abs(back - r).max() <= 2 * tol(dtype, n, r, 1.0 / n), n
        spec = pairs(ref, dtype)
        c2r = to_host(F.irfft(on(spec, device), n=n))
        assert np.abs(c2r - np.fft.irfft(ref, n=n, axis=-1)).max() <= 2 * tol(dtype, n, spec, 1.0 / n), n


@pytest.mark.parametrize("device", DEVICES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_rfftn_irfftn_match_numpy(dev, device, dtype):
    r = np.random.default_rng(3).uniform(-1, 1, (4, 5, 6)).astype(dtype)
    for axes in ((0, 1, 2), (1, 2), (2, 0)):
        got = to_host(F.rfftn(on(r, device), axes=axes))
        ref = np.fft.rfftn(r.astype(np.float64), axes=axes)
        pts = int(np.prod([r.shape[a] for a in axes]))
        assert got.shape == ref.shape + (2,), axes
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, pts, r, 1.0), axes
        s = [r.shape[a] for a in axes]
        back = to_host(F.irfftn(on(got, device), s=s, axes=axes))
        assert np.abs(back - r).max() <= 2 * tol(dtype, pts, r, 1.0 / pts), axes
    # the input of a complex-to-real transform is left unchanged
    spec = pairs(np.fft.rfftn(r.astype(np.float64)), dtype)
    keep = spec.copy()
    sd = on(spec, device)
    F.irfftn(sd, s=r.shape, axes=(0, 1, 2))
    assert np.array_equal(to_host(sd), keep)


def test_plans_are_cached_and_carry_their_shapes(dev):
    a = F.plan("r2c", "float32", (4, 9), (1,))
    assert a is F.plan("r2c", "float32", (4, 9), (-1,))
    assert a.input_shape == (4, 9) and a.output_shape == (4, 5, 2)
    assert F.plan("c2c", "float64", (3, 4), (0, 1), forward=False).input_shape == (3, 4, 2)
    assert F.plan("c2c", "float32", (8,), (0,), device="cuda") is not F.plan("c2c", "float32", (8,), (0,))


def test_misuse_is_a_named_error(dev):
    p = F.plan("c2c", "float32", (8,), (0,))
    with pytest.raises(aether.Error, match="plan wants float32"):
        p.execute(aether.zeros((8, 2), "float64"), aether.zeros((8, 2), "float64"))
    with pytest.raises(aether.Error, match="wrong shape"):
        p.execute(aether.zeros((7, 2), "float32"), aether.zeros((8, 2), "float32"))
    with pytest.raises(aether.Error, match="plan's device"):
        p.execute(aether.zeros((8, 2), "float32", "cuda"), aether.zeros((8, 2), "float32"))
    pd = F.plan("r2c", "float32", (8,), (0,), device="cuda")
    x = aether.zeros((8,), "float32", "cuda")
    with pytest.raises(aether.Error, match="wrong shape"):
        pd.execute(x, x)
    with pytest.raises(F.Error):
        F.plan("c2c", "float32", ((1 << 25) + 1,), (0,), device="cuda")
    with pytest.raises(aether.FftUnavailable):
        F.plan("c2c", "float64", (1 << 27,), (0,), device="cuda")
    # pinned host memory is host memory: it runs the host path and gives a host array
    pin = aether.pinned((8, 2), "float32")
    pin.zero()
    assert F.fft(pin).device.kind == "cpu"


def test_the_host_path_needs_no_numpy(dev):
    code = textwrap.dedent("""
        import sys
        for m in ('numpy', 'cupy', 'torch'):
            sys.modules[m] = None
        import aether_amd.dsc as aether
        from aether_amd.dsc import fft
        x = aether.asarray([[1.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0]])
        y = fft.fft(x)
        assert [list(r) for r in memoryview(y).tolist()] == [[1.0, 0.0]] * 4, memoryview(y).tolist()
        print('ok')
    """)
    import os

    env = dict(os.environ, PYTHONPATH=str(ROOT))
    out = subprocess.run([sys.executable, "-W", "ignore", "-c", code], capture_output=True, text=True, timeout=120,
                         env=env)
    assert out.stdout.strip() == "ok", out.stderr[-600:]


def test_a_plan_built_before_a_capture_runs_inside_it_and_replays(dev):
    s = torch.cuda.Stream()
    x = rand_c((16, 64), "float32", 1)
    d = aether.asarray(x, device="cuda")
    out = aether.empty((16, 64, 2), "float32", "cuda", stream=s)
    p = F.plan("c2c", "float32", (16, 64), (1,), True, "backward", "cuda", s)  # eager, outside the capture
    p2 = F.plan("c2r", "float32", (16, 64), (1,), False, "backward", "cuda", s)
    real = aether.empty((16, 64), "float32", "cuda", stream=s)
    ref = np.fft.fft(cplx(x), axis=-1)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.execute(d, out, s)
    tout = torch.from_dlpack(out)
    for _ in range(2):
        tout.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.abs(cplx(tout.cpu().numpy()) - ref).max() <= tol("float32", 64, x, 1.0)
    # a complex-to-real plan (work copy, Hermitian extension, real part) also captures and replays
    r = np.random.default_rng(2).uniform(-1, 1, (16, 64)).astype("float32")
    sp = pairs(np.fft.rfft(r.astype(np.float64), axis=-1), "float32")
    dsp = aether.asarray(sp, device="cuda")
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=s):
        p2.execute(dsp, real, s)
    treal = torch.from_dlpack(real)
    for _ in range(2):
        treal.zero_()
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.abs(treal.cpu().numpy() - r).max() <= 2 * tol("float32", 64, r, 1 / 64)


def test_plan_creation_inside_a_capture_is_refused_by_name(dev):
    s = torch.cuda.Stream()
    y = torch.zeros(4, device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        y += 1
        with pytest.raises(aether.Error, match="plan creation is not allowed inside a stream capture"):
            F.plan("c2c", "float32", (123, 45), (1,), True, "backward", "cuda", s)
    g.replay()
    torch.cuda.synchronize()
    assert float(y.sum()) == 4.0


@pytest.mark.parametrize("device", DEVICES)
def test_repeated_transforms_give_the_same_bits(dev, device):
    x = rand_c((64, 256), "float32", 11)
    a = to_host(F.fft(on(x, device)))
    b = to_host(F.fft(on(x, device)))
    assert a.tobytes() == b.tobytes()
    r = np.random.default_rng(12).uniform(-1, 1, (64, 100)).astype("float64")
    assert to_host(F.rfftn(on(r, device))).tobytes() == to_host(F.rfftn(on(r, device))).tobytes()
