"""GPU: aether_amd.fft -- the reference's aether.fft API for torch tensors -- against numpy.fft with the
reference's error bound (tests/fft_rows.py). Beyond the rows of the array package's test
(test_gpu_dsc_fft.py): lengths past the radix kernels' 8192 (four-step for powers of two, two levels of it
past 2^26, Bluestein's chirp-z with a four-step convolution otherwise), a long axis that is not the last one, long real
transforms, host tensors coming back on the host, and a Bluestein plan and a c2r plan captured into one
graph and replayed."""
import math

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import fft as F
from fft_rows import LENGTHS, MULTI_AXES, REAL_LENGTHS, Data, bound, scale, to_complex, to_pairs, worst

pytestmark = pytest.mark.gpu


def gpu(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def back(t):
    return t.cpu().numpy()


def check_c2c(n, lines, dtype):
    x = Data(n).complex_pairs((lines, n), dtype)
    for forward, transform, numpy_ref in ((True, F.fft, np.fft.fft), (False, F.ifft, np.fft.ifft)):
        err = worst(to_complex(back(transform(gpu(x)))), numpy_ref(to_complex(x), axis=-1))
        assert err <= bound(dtype, n, x, scale("backward", forward, n)), (n, forward, err)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_c2c_lengths(dev, dtype):
    for n in LENGTHS + (4095, 4096, 8192):
        check_c2c(n, 3, dtype)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("n,lines", [(16384, 3), (32768, 2), (1 << 16, 1), (1 << 20, 1), (5000, 2), (10007, 1),
                                     (65537, 1), (300000, 1)])
def test_long_lengths(dev, dtype, n, lines):
    check_c2c(n, lines, dtype)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_long_axes_inside_larger_arrays(dev, dtype):
    data = Data(21)
    x = data.complex_pairs((16384, 3), dtype)  # lines strided by the inner extent
    assert worst(to_complex(back(F.fft(gpu(x), axis=0))), np.fft.fft(to_complex(x), axis=0)) <= \
        bound(dtype, 16384, x, 1.0)
    y = data.complex_pairs((4, 20000), dtype)
    assert worst(to_complex(back(F.fftn(gpu(y), axes=(0, 1)))), np.fft.fftn(to_complex(y), axes=(0, 1))) <= \
        bound(dtype, 80000, y, 1.0)
    for n in (1 << 15, 50000):
        r = data.real((2, n), dtype)
        spec = back(F.rfft(gpu(r)))
        assert worst(to_complex(spec), np.fft.rfft(r.astype(np.float64), axis=-1)) <= bound(dtype, n, r, 1.0), n
        assert worst(back(F.irfft(gpu(spec), n=n)), r) <= 2 * bound(dtype, n, spec, 1.0 / n), n


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("axes", MULTI_AXES, ids=str)
def test_multi_axis_round_trip(dev, dtype, axes):
    x = Data(5).complex_pairs((4, 7, 6), dtype)
    pts = math.prod(x.shape[a] for a in axes)
    spec = back(F.fftn(gpu(x), axes=axes))
    assert worst(to_complex(spec), np.fft.fftn(to_complex(x), axes=axes)) <= bound(dtype, pts, x, 1.0)
    assert worst(to_complex(back(F.ifftn(gpu(spec), axes=axes))), to_complex(x)) <= 2 * bound(dtype, pts, spec, 1 / pts)


@pytest.mark.parametrize("norm", ["backward", "forward", "ortho"])
def test_norms(dev, norm):
    x = Data(9).complex_pairs((2, 12), "float64")
    for forward, transform, numpy_ref in ((True, F.fft, np.fft.fft), (False, F.ifft, np.fft.ifft)):
        err = worst(to_complex(back(transform(gpu(x), norm=norm))), numpy_ref(to_complex(x), norm=norm))
        assert err <= bound("float64", 12, x, scale(norm, forward, 12))


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_real_one_axis(dev, dtype):
    for n in REAL_LENGTHS:
        r = Data(n).real((3, n), dtype)
        ref = np.fft.rfft(r.astype(np.float64), axis=-1)
        spec = back(F.rfft(gpu(r)))
        assert spec.shape == ref.shape + (2,) and worst(to_complex(spec), ref) <= bound(dtype, n, r, 1.0), n
        assert worst(back(F.irfft(gpu(spec), n=n)), r) <= 2 * bound(dtype, n, r, 1.0 / n), n
        numpy_spec = to_pairs(ref, dtype)
        assert worst(back(F.irfft(gpu(numpy_spec), n=n)), np.fft.irfft(ref, n=n, axis=-1)) <= \
            2 * bound(dtype, n, numpy_spec, 1.0 / n), n


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_real_long_lengths_normwise(dev, dtype):
    """Bluestein real lengths: forward within the bound; the round trip within the normwise bound of the
    inverse, whose input (the spectrum) has norm sqrt(n) ||x||_2, so the scale 1/n gives ||x||_2 / sqrt(n)."""
    for n in (1000, 2049, 4095):
        r = Data(n).real((3, n), dtype)
        spec = back(F.rfft(gpu(r)))
        assert worst(to_complex(spec), np.fft.rfft(r.astype(np.float64), axis=-1)) <= bound(dtype, n, r, 1.0), n
        assert worst(back(F.irfft(gpu(spec), n=n)), r) <= 2 * bound(dtype, n, r, 1.0 / math.sqrt(n)), n


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_real_several_axes(dev, dtype):
    r = Data(3).real((4, 5, 6), dtype)
    for axes in ((0, 1, 2), (1, 2), (2, 0)):
        pts = math.prod(r.shape[a] for a in axes)
        ref = np.fft.rfftn(r.astype(np.float64), axes=axes)
        spec = back(F.rfftn(gpu(r), axes=axes))
        assert spec.shape == ref.shape + (2,) and worst(to_complex(spec), ref) <= bound(dtype, pts, r, 1.0), axes
        inv = back(F.irfftn(gpu(spec), s=[r.shape[a] for a in axes], axes=axes))
        assert worst(inv, r) <= 2 * bound(dtype, pts, r, 1.0 / pts), axes
    src = gpu(to_pairs(np.fft.rfftn(r.astype(np.float64)), dtype))
    kept = src.clone()
    F.irfftn(src, s=r.shape, axes=(0, 1, 2))
    torch.cuda.synchronize()
    assert torch.equal(src, kept)  # the c2r input is left as it was


def test_host_inputs_give_host_outputs(dev):
    x = Data(4).complex_pairs((2, 16), "float64")
    y = F.fft(torch.from_numpy(x))
    assert not y.is_cuda
    assert worst(to_complex(y.numpy()), np.fft.fft(to_complex(x), axis=-1)) <= bound("float64", 16, x, 1.0)
    assert np.array_equal(F.fft(x).numpy(), y.numpy())  # a numpy array in


def test_plans(dev):
    p = F.plan("r2c", "float32", (4, 9), (1,))
    assert F.plan("r2c", "float32", (4, 9), (-1,)) is p and (p.input_shape, p.output_shape) == ((4, 9), (4, 5, 2))
    assert F.plan("c2c", "float64", (3, 4), (0, 1), forward=False).input_shape == (3, 4, 2)
    assert F.plan("c2r", "float64", (3, 8), (1,)).input_shape == (3, 5, 2)


def test_named_misuse(dev):
    x = gpu(Data(0).complex_pairs((8,), "float32"))
    p = F.plan("c2c", "float32", (8,), (0,))
    f64 = torch.zeros((8, 2), dtype=torch.float64, device="cuda")
    cases = [(lambda: F.fft(torch.zeros(8, dtype=torch.float32, device="cuda")), "trailing axis of 2"),
             (lambda: F.fft(x, norm="unitary"), "norm"), (lambda: F.fft(x, axis=3), "out of range"),
             (lambda: F.plan("c2c", "int32", (4,), (0,)), "float32 or float64"),
             (lambda: p.execute(f64, f64), "plan wants float32"),
             (lambda: p.execute(torch.zeros((7, 2), device="cuda"), torch.zeros((8, 2), device="cuda")), "wrong shape")]
    for call, words in cases:
        with pytest.raises(ae.AetherError, match=words):
            call()
    for shape in (((1 << 32) + 1,), (1 << 35,)):  # Bluestein stops below 2^32, the four-step at 2^34
        with pytest.raises(F.FftUnavailable):
            F.plan("c2c", "float64", shape, (0,))


@pytest.mark.parametrize("n", [1 << 27, (1 << 25) + 1], ids=["pow2_2^27", "bluestein_2^25+1"])
def test_lengths_past_two_levels(dev, n):
    """A power of two past 2^26 (columns of 8192 points, rows of 2^14 points that are a four-step pass
    themselves) and a Bluestein length whose convolution is such a transform (m = 2^27), forward and
    inverse, against numpy.fft with the reference's bound; the round trip within twice the bound."""
    x = Data(27).complex_pairs((n,), "float64")
    xt = gpu(x)
    X = F.fft(xt)
    err = worst(to_complex(back(X)), np.fft.fft(to_complex(x)))
    assert err <= bound("float64", n, x, 1.0), err
    xb = F.ifft(X)
    assert worst(to_complex(back(xb)), to_complex(x)) <= 2 * bound("float64", n, x, 1.0 / n) * n


def test_bluestein_and_c2r_plans_replay_in_one_graph(dev):
    side = torch.cuda.Stream()
    data = Data(1)
    x = data.complex_pairs((16, 100), "float32")  # a Bluestein length: five kernels per pass in the graph
    r = data.real((16, 64), "float32")
    c2c = F.plan("c2c", "float32", (16, 100), (1,), True, "backward", "cuda", side)
    c2r = F.plan("c2r", "float32", (16, 64), (1,), False, "backward", "cuda", side)
    src, half = gpu(x), gpu(to_pairs(np.fft.rfft(r.astype(np.float64), axis=-1), "float32"))
    spec = torch.empty((16, 100, 2), dtype=torch.float32, device="cuda")
    real = torch.empty((16, 64), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        c2c.execute(src, spec, side)
        c2r.execute(half, real, side)
    want = np.fft.fft(to_complex(x), axis=-1)
    for _ in range(3):
        spec.zero_()
        real.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert worst(to_complex(back(spec)), want) <= bound("float32", 100, x, 1.0)
        assert worst(back(real), r) <= 2 * bound("float32", 64, r, 1 / 64)
