"""GPU: aether_amd.fft, the reference's Python FFT API (dsc/aether/fft.py) over this library's seam plans,
held to the reference's own device tests (dsc/tests/test_fft.py:78-160, 206-268): numpy.fft is the oracle and
the tolerance is ULP-scaled, (8 + 4 log2 N) * eps * ||x||_2 * scale for a transform over N points.
Non-power-of-two lengths run Bluestein's chirp-z; both widths, every norm, real transforms, plans and
their errors, and a plan built before a graph capture replaying inside it."""
import math

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import fft as F

pytestmark = pytest.mark.gpu

DTYPES = ["float32", "float64"]
NP = {"float32": np.float32, "float64": np.float64}


def pairs(c, dtype):
    return np.ascontiguousarray(np.stack([c.real, c.imag], -1).astype(NP[dtype]))


def cplx(p):
    p = np.asarray(p, dtype=np.float64)
    return p[..., 0] + 1j * p[..., 1]


def tol(dtype, points, x, scale):
    return (8 + 4 * math.log2(points + 1)) * np.finfo(NP[dtype]).eps * float(np.linalg.norm(np.asarray(x, np.float64))) * scale


def rand_c(shape, dtype, seed=0):
    r = np.random.default_rng(seed)
    return pairs(r.uniform(-1, 1, shape) + 1j * r.uniform(-1, 1, shape), dtype)


def scale_of(norm, forward, n):
    if norm == "ortho":
        return 1 / math.sqrt(n)
    return 1 / n if forward == (norm == "forward") else 1.0


def on(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("dtype", DTYPES)
def test_c2c_lengths_match_numpy(dev, dtype):
    for n in (1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 17, 31, 64, 100, 127, 1000, 4096, 4095, 8192):
        x = rand_c((3, n), dtype, n)
        for forward in (True, False):
            fn = F.fft if forward else F.ifft
            got = host(fn(on(x)))
            ref = (np.fft.fft if forward else np.fft.ifft)(cplx(x), axis=-1)
            err = np.abs(cplx(got) - ref).max()
            assert err <= tol(dtype, n, x, scale_of("backward", forward, n)), (n, forward, err)


@pytest.mark.parametrize("dtype", DTYPES)
def test_long_lengths_match_numpy(dev, dtype):
    """Powers of two above 8192 (four-step: columns, twiddle, rows, transposed read-out) and other lengths
    whose Bluestein convolution is longer than 8192 (its convolution then runs four-step), both directions,
    held to the same ULP-scaled bound."""
    for n, lines in ((16384, 3), (32768, 2), (1 << 16, 1), (1 << 20, 1), (5000, 2), (10007, 1), (65537, 1),
                     (300000, 1)):
        x = rand_c((lines, n), dtype, n)
        for forward in (True, False):
            fn = F.fft if forward else F.ifft
            got = host(fn(on(x)))
            ref = (np.fft.fft if forward else np.fft.ifft)(cplx(x), axis=-1)
            err = np.abs(cplx(got) - ref).max()
            assert err <= tol(dtype, n, x, scale_of("backward", forward, n)), (n, forward, err)


@pytest.mark.parametrize("dtype", DTYPES)
def test_long_axis_strided_and_inside_nd(dev, dtype):
    """A long axis that is not the last one (its lines are strided by the inner extent), a long axis in a
    multi-axis transform, and real transforms over long axes."""
    x = rand_c((16384, 3), dtype, 21)
    got = host(F.fft(on(x), axis=0))
    ref = np.fft.fft(cplx(x), axis=0)
    assert np.abs(cplx(got) - ref).max() <= tol(dtype, 16384, x, 1.0)
    x = rand_c((4, 20000), dtype, 22)
    got = host(F.fftn(on(x), axes=(0, 1)))
    ref = np.fft.fftn(cplx(x), axes=(0, 1))
    assert np.abs(cplx(got) - ref).max() <= tol(dtype, 80000, x, 1.0)
    for n in (1 << 15, 50000):
        r = np.random.default_rng(n).uniform(-1, 1, (2, n)).astype(NP[dtype])
        spec = host(F.rfft(on(r)))
        ref = np.fft.rfft(r.astype(np.float64), axis=-1)
        assert np.abs(cplx(spec) - ref).max() <= tol(dtype, n, r, 1.0), n
        back = host(F.irfft(on(spec), n=n))
        assert np.abs(back - r).max() <= 2 * tol(dtype, n, spec, 1.0 / n), n


@pytest.mark.parametrize("dtype", DTYPES)
def test_multi_axis_and_batched_match_numpy(dev, dtype):
    x = rand_c((4, 7, 6), dtype, 5)
    xc = cplx(x)
    cases = [((1,), np.fft.fft(xc, axis=1)), ((0, 2), np.fft.fftn(xc, axes=(0, 2))), ((0, 1, 2), np.fft.fftn(xc)),
             ((2, 0), np.fft.fftn(xc, axes=(2, 0)))]
    for axes, ref in cases:
        got = host(F.fftn(on(x), axes=axes))
        pts = int(np.prod([x.shape[a] for a in axes]))
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, pts, x, 1.0), axes
        back = host(F.ifftn(on(got), axes=axes))
        assert np.abs(cplx(back) - xc).max() <= 2 * tol(dtype, pts, got, 1.0 / pts), axes


@pytest.mark.parametrize("norm", ["backward", "forward", "ortho"])
def test_norms_match_numpy(dev, norm):
    x = rand_c((2, 12), "float64", 9)
    for forward in (True, False):
        fn = F.fft if forward else F.ifft
        got = host(fn(on(x), norm=norm))
        ref = (np.fft.fft if forward else np.fft.ifft)(cplx(x), norm=norm)
        assert np.abs(cplx(got) - ref).max() <= tol("float64", 12, x, scale_of(norm, forward, 12))


@pytest.mark.parametrize("dtype", DTYPES)
def test_rfft_irfft_match_numpy(dev, dtype):
    for n in (2, 3, 7, 8, 9, 16, 17, 50, 128):
        r = np.random.default_rng(n).uniform(-1, 1, (3, n)).astype(NP[dtype])
        got = host(F.rfft(on(r)))
        ref = np.fft.rfft(r.astype(np.float64), axis=-1)
        assert got.shape == ref.shape + (2,)
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, n, r, 1.0), n
        back = host(F.irfft(on(got), n=n))
        assert np.abs(back - r).max() <= 2 * tol(dtype, n, r, 1.0 / n), n
        spec = pairs(ref, dtype)
        c2r = host(F.irfft(on(spec), n=n))
        assert np.abs(c2r - np.fft.irfft(ref, n=n, axis=-1)).max() <= 2 * tol(dtype, n, spec, 1.0 / n), n


@pytest.mark.parametrize("dtype", DTYPES)
def test_long_real_lengths_normwise(dev, dtype):
    """Bluestein lengths beyond the reference's list: the forward transform within the reference's bound;
    the round trip within the normwise bound of the inverse, whose input (the spectrum) has norm
    sqrt(n) ||x||_2, so scale 1/n gives ||x||_2 / sqrt(n)."""
    for n in (1000, 2049, 4095):
        r = np.random.default_rng(n).uniform(-1, 1, (3, n)).astype(NP[dtype])
        got = host(F.rfft(on(r)))
        ref = np.fft.rfft(r.astype(np.float64), axis=-1)
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, n, r, 1.0), n
        back = host(F.irfft(on(got), n=n))
        assert np.abs(back - r).max() <= 2 * tol(dtype, n, r, 1.0 / math.sqrt(n)), n


@pytest.mark.parametrize("dtype", DTYPES)
def test_rfftn_irfftn_match_numpy(dev, dtype):
    r = np.random.default_rng(3).uniform(-1, 1, (4, 5, 6)).astype(NP[dtype])
    for axes in ((0, 1, 2), (1, 2), (2, 0)):
        got = host(F.rfftn(on(r), axes=axes))
        ref = np.fft.rfftn(r.astype(np.float64), axes=axes)
        pts = int(np.prod([r.shape[a] for a in axes]))
        assert got.shape == ref.shape + (2,), axes
        assert np.abs(cplx(got) - ref).max() <= tol(dtype, pts, r, 1.0), axes
        s = [r.shape[a] for a in axes]
        back = host(F.irfftn(on(got), s=s, axes=axes))
        assert np.abs(back - r).max() <= 2 * tol(dtype, pts, r, 1.0 / pts), axes
    # the input of a complex-to-real transform is left unchanged
    spec = on(pairs(np.fft.rfftn(r.astype(np.float64)), dtype))
    keep = spec.clone()
    F.irfftn(spec, s=r.shape, axes=(0, 1, 2))
    torch.cuda.synchronize()
    assert torch.equal(spec, keep)


def test_host_arrays_come_back_on_the_host(dev):
    x = rand_c((2, 16), "float64", 4)
    y = F.fft(torch.from_numpy(x))
    assert not y.is_cuda
    assert np.abs(cplx(y.numpy()) - np.fft.fft(cplx(x), axis=-1)).max() <= tol("float64", 16, x, 1.0)
    z = F.fft(x)  # a numpy array
    assert np.array_equal(z.numpy(), y.numpy())


def test_plans_are_cached_and_carry_their_shapes(dev):
    a = F.plan("r2c", "float32", (4, 9), (1,))
    assert a is F.plan("r2c", "float32", (4, 9), (-1,))
    assert a.input_shape == (4, 9) and a.output_shape == (4, 5, 2)
    assert F.plan("c2c", "float64", (3, 4), (0, 1), forward=False).input_shape == (3, 4, 2)
    assert F.plan("c2r", "float64", (3, 8), (1,)).input_shape == (3, 5, 2)


def test_misuse_is_a_named_error(dev):
    x = on(rand_c((8,), "float32"))
    with pytest.raises(ae.AetherError, match="trailing axis of 2"):
        F.fft(torch.zeros(8, dtype=torch.float32, device="cuda"))
    with pytest.raises(ae.AetherError, match="norm"):
        F.fft(x, norm="bogus")
    with pytest.raises(ae.AetherError, match="out of range"):
        F.fft(x, axis=3)
    with pytest.raises(ae.AetherError, match="float32 or float64"):
        F.plan("c2c", "int32", (4,), (0,))
    p = F.plan("c2c", "float32", (8,), (0,))
    with pytest.raises(ae.AetherError, match="plan wants float32"):
        p.execute(torch.zeros((8, 2), dtype=torch.float64, device="cuda"),
                  torch.zeros((8, 2), dtype=torch.float64, device="cuda"))
    with pytest.raises(ae.AetherError, match="wrong shape"):
        p.execute(torch.zeros((7, 2), dtype=torch.float32, device="cuda"),
                  torch.zeros((8, 2), dtype=torch.float32, device="cuda"))
    with pytest.raises(F.FftUnavailable):
        F.plan("c2c", "float32", ((1 << 25) + 1,), (0,))  # not a power of two and beyond 2^25
    with pytest.raises(F.FftUnavailable):
        F.plan("c2c", "float64", (1 << 27,), (0,))  # a power of two beyond 2^26


def test_a_plan_built_before_a_capture_runs_inside_it_and_replays(dev):
    s = torch.cuda.Stream()
    x = rand_c((16, 100), "float32", 1)  # a Bluestein length: five kernels per pass inside the graph
    d = on(x)
    out = torch.empty((16, 100, 2), dtype=torch.float32, device="cuda")
    p = F.plan("c2c", "float32", (16, 100), (1,), True, "backward", "cuda", s)
    r = np.random.default_rng(2).uniform(-1, 1, (16, 64)).astype(np.float32)
    p2 = F.plan("c2r", "float32", (16, 64), (1,), False, "backward", "cuda", s)
    sp = on(pairs(np.fft.rfft(r.astype(np.float64), axis=-1), "float32"))
    real = torch.empty((16, 64), dtype=torch.float32, device="cuda")
    ref = np.fft.fft(cplx(x), axis=-1)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.execute(d, out, s)
        p2.execute(sp, real, s)
    for _ in range(3):
        out.zero_()
        real.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.abs(cplx(host(out)) - ref).max() <= tol("float32", 100, x, 1.0)
        assert np.abs(host(real) - r).max() <= 2 * tol("float32", 64, r, 1 / 64)
