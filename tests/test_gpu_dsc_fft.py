"""GPU: aether_amd.dsc.fft over aether arrays, on host arrays (staged through the device) and on device
arrays.

The acceptance rows are the reference package's FFT rows (dsc/tests/test_fft.py), restated: c2c over many
lengths (powers of two, primes, composites), batched and multi-axis transforms and their round trips, the
three norms, r2c / c2r including a spectrum built by numpy, plan caching and plan shapes, named misuse
errors, the host path without numpy, a plan made before a capture executed inside it and replayed, plan
creation refused inside a capture, and bit-identical repeats. numpy.fft in complex128 is the reference and
the bound is the reference's: max |error| <= (8 + 4 log2(N + 1)) * eps * ||x||_2 * s, with s the
transform's scale (1, 1/N or 1/sqrt(N))."""
import math
import os
import pathlib
import subprocess
import sys
import textwrap

import numpy as np
import pytest
import torch

import aether_amd.dsc as aether
from aether_amd.dsc import fft as F
from fft_rows import LENGTHS, MULTI_AXES, REAL_LENGTHS, Data, bound, scale, to_complex, to_pairs, worst

pytestmark = pytest.mark.gpu
REPO = pathlib.Path(__file__).resolve().parents[1]
PLACES = ["cpu", "cuda"]


def fetch(a):
    """A host numpy copy of an aether array wherever it lives."""
    return np.array(aether.asarray(a, device="cpu"), copy=True)


@pytest.mark.parametrize("place", PLACES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("forward", [True, False], ids=["fft", "ifft"])
def test_c2c_over_many_lengths(dev, place, dtype, forward):
    transform, numpy_ref = (F.fft, np.fft.fft) if forward else (F.ifft, np.fft.ifft)
    for n in LENGTHS:
        x = Data(1000 + n).complex_pairs((3, n), dtype)
        y = transform(aether.asarray(x, device=place))
        assert y.device.kind == place and y.shape == (3, n, 2)
        err = worst(to_complex(fetch(y)), numpy_ref(to_complex(x), axis=-1))
        assert err <= bound(dtype, n, x, scale("backward", forward, n)), (n, err)


@pytest.mark.parametrize("place", PLACES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("axes", MULTI_AXES, ids=str)
def test_batched_and_multi_axis_round_trip(dev, place, dtype, axes):
    x = Data(77).complex_pairs((4, 7, 6), dtype)
    pts = math.prod(x.shape[a] for a in axes)
    spectrum = F.fftn(aether.asarray(x, device=place), axes=axes)
    got = fetch(spectrum)
    assert worst(to_complex(got), np.fft.fftn(to_complex(x), axes=axes)) <= bound(dtype, pts, x, 1.0)
    back = fetch(F.ifftn(aether.asarray(got, device=place), axes=axes))
    assert worst(to_complex(back), to_complex(x)) <= 2 * bound(dtype, pts, got, 1.0 / pts)


@pytest.mark.parametrize("place", PLACES)
@pytest.mark.parametrize("norm", ["backward", "forward", "ortho"])
def test_norms(dev, place, norm):
    x = Data(9).complex_pairs((2, 12), "float64")
    for forward, transform, numpy_ref in ((True, F.fft, np.fft.fft), (False, F.ifft, np.fft.ifft)):
        got = to_complex(fetch(transform(aether.asarray(x, device=place), norm=norm)))
        assert worst(got, numpy_ref(to_complex(x), norm=norm)) <= bound("float64", 12, x, scale(norm, forward, 12))


@pytest.mark.parametrize("place", PLACES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("n", REAL_LENGTHS)
def test_real_transforms_one_axis(dev, place, dtype, n):
    r = Data(n).real((3, n), dtype)
    ref = np.fft.rfft(r.astype(np.float64), axis=-1)
    spec = fetch(F.rfft(aether.asarray(r, device=place)))
    assert spec.shape == (3, n // 2 + 1, 2)
    assert worst(to_complex(spec), ref) <= bound(dtype, n, r, 1.0)
    back = fetch(F.irfft(aether.asarray(spec, device=place), n=n))
    assert worst(back, r) <= 2 * bound(dtype, n, r, 1.0 / n)
    # a spectrum made by numpy, inverted by the plan
    numpy_spec = to_pairs(ref, dtype)
    inv = fetch(F.irfft(aether.asarray(numpy_spec, device=place), n=n))
    assert worst(inv, np.fft.irfft(ref, n=n, axis=-1)) <= 2 * bound(dtype, n, numpy_spec, 1.0 / n)


@pytest.mark.parametrize("place", PLACES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_real_transforms_several_axes(dev, place, dtype):
    r = Data(3).real((4, 5, 6), dtype)
    for axes in ((0, 1, 2), (1, 2), (2, 0)):
        pts = math.prod(r.shape[a] for a in axes)
        ref = np.fft.rfftn(r.astype(np.float64), axes=axes)
        spec = fetch(F.rfftn(aether.asarray(r, device=place), axes=axes))
        assert spec.shape == ref.shape + (2,) and worst(to_complex(spec), ref) <= bound(dtype, pts, r, 1.0)
        lengths = [r.shape[a] for a in axes]
        back = fetch(F.irfftn(aether.asarray(spec, device=place), s=lengths, axes=axes))
        assert worst(back, r) <= 2 * bound(dtype, pts, r, 1.0 / pts)
    # the complex-to-real transform leaves its input as it was
    spec = to_pairs(np.fft.rfftn(r.astype(np.float64)), dtype)
    src = aether.asarray(spec.copy(), device=place)
    F.irfftn(src, s=r.shape, axes=(0, 1, 2))
    assert np.array_equal(fetch(src), spec)


def test_plan_cache_and_shapes(dev):
    p = F.plan("r2c", "float32", (4, 9), (1,))
    assert F.plan("r2c", "float32", (4, 9), (-1,)) is p
    assert (p.input_shape, p.output_shape) == ((4, 9), (4, 5, 2))
    assert F.plan("c2c", "float64", (3, 4), (0, 1), forward=False).input_shape == (3, 4, 2)
    assert F.plan("c2c", "float32", (8,), (0,), device="cuda") is not F.plan("c2c", "float32", (8,), (0,))


def test_misuse_is_named(dev):
    host_plan = F.plan("c2c", "float32", (8,), (0,))
    wrong = [
        (lambda: host_plan.execute(aether.zeros((8, 2), "float64"), aether.zeros((8, 2), "float64")),
         "plan wants float32"),
        (lambda: host_plan.execute(aether.zeros((7, 2), "float32"), aether.zeros((8, 2), "float32")), "wrong shape"),
        (lambda: host_plan.execute(aether.zeros((8, 2), "float32", "cuda"), aether.zeros((8, 2), "float32")),
         "plan's device"),
    ]
    dev_plan = F.plan("r2c", "float32", (8,), (0,), device="cuda")
    real = aether.zeros(8, "float32", "cuda")
    wrong.append((lambda: dev_plan.execute(real, real), "wrong shape"))
    for call, words in wrong:
        with pytest.raises(aether.Error, match=words):
            call()
    with pytest.raises(aether.Error):  # neither a power of two nor below Bluestein's 2^32
        F.plan("c2c", "float32", ((1 << 32) + 1,), (0,), device="cuda")
    with pytest.raises(aether.FftUnavailable):  # a power of two beyond 2^34
        F.plan("c2c", "float64", (1 << 35,), (0,), device="cuda")
    pin = aether.pinned((8, 2), "float32")
    pin.zero()
    assert F.fft(pin).device.kind == "cpu"  # pinned is host memory: host path, host result


def test_host_path_without_numpy(dev):
    source = textwrap.dedent("""
        import sys
        sys.modules.update(numpy=None, cupy=None, torch=None)
        import aether_amd.dsc as aether
        impulse = aether.asarray([[1.0, 0.0]] + [[0.0, 0.0]] * 3)
        flat = aether.fft.fft(impulse)
        print(memoryview(flat).tolist() == [[1.0, 0.0]] * 4)
    """)
    child = subprocess.run([sys.executable, "-W", "ignore", "-c", source], capture_output=True, text=True,
                           timeout=120, env={**os.environ, "PYTHONPATH": str(REPO)})
    assert child.stdout.strip() == "True", child.stderr[-800:]


def replay_and_check(graph, target, expect, tol, times=2):
    view = torch.from_dlpack(target)
    for _ in range(times):
        view.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert worst(expect[0](view.cpu().numpy()), expect[1]) <= tol


def test_plans_made_before_a_capture_replay(dev):
    side = torch.cuda.Stream()
    data = Data(1)
    x = data.complex_pairs((16, 64), "float32")
    c2c = F.plan("c2c", "float32", (16, 64), (1,), True, "backward", "cuda", side)
    c2r = F.plan("c2r", "float32", (16, 64), (1,), False, "backward", "cuda", side)
    src = aether.asarray(x, device="cuda")
    spec = aether.empty((16, 64, 2), "float32", "cuda", stream=side)
    r = data.real((16, 64), "float32")
    half = aether.asarray(to_pairs(np.fft.rfft(r.astype(np.float64), axis=-1), "float32"), device="cuda")
    real = aether.empty((16, 64), "float32", "cuda", stream=side)
    side.synchronize()
    g1, g2 = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1, stream=side):
        c2c.execute(src, spec, side)
    with torch.cuda.graph(g2, stream=side):
        c2r.execute(half, real, side)
    replay_and_check(g1, spec, (to_complex, np.fft.fft(to_complex(x), axis=-1)), bound("float32", 64, x, 1.0))
    replay_and_check(g2, real, (lambda v: v, r), 2 * bound("float32", 64, r, 1.0 / 64))


def test_no_plan_is_made_inside_a_capture(dev):
    side = torch.cuda.Stream()
    counter = torch.zeros(4, device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        counter += 1
        with pytest.raises(aether.Error, match="plan creation is not allowed inside a stream capture"):
            F.plan("c2c", "float32", (123, 45), (1,), True, "backward", "cuda", side)
    g.replay()
    torch.cuda.synchronize()
    assert counter.sum().item() == 4.0


@pytest.mark.parametrize("place", PLACES)
def test_repeats_are_bit_identical(dev, place):
    data = Data(11)
    x = aether.asarray(data.complex_pairs((64, 256), "float32"), device=place)
    assert fetch(F.fft(x)).tobytes() == fetch(F.fft(x)).tobytes()
    r = aether.asarray(data.real((64, 100), "float64"), device=place)
    assert fetch(F.rfftn(r)).tobytes() == fetch(F.rfftn(r)).tobytes()
