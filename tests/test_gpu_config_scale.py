"""GPU parity at the workloads' full sizes (the sizes the bench reports), against the oracle on windows of
the full outputs and through size-independent properties:

* Philox4x32-10, seed 42 (key = seed, counter = sample index): the full 2^30-word, 2^29-uniform and
  2^28-normal streams generated in one launch each; windows at the head, in the middle and at the end
  equal the oracle's stream at the same positions bit for bit (tests/test_Random.cpp:356-407 asks the
  streams bit-exact; the oracle's normals use the device's own math, so they are bit-exact too).
* The fused FP64 expression and the Vec3 expression at 2^27 samples: windows bit for bit against the
  oracle run on the same input windows; Vec3 everywhere within 1 ulp of the same folds spelled as
  separate torch operations.
* The FP32 form at 2^28 samples: d within a few float ulps of the FP64 evaluation of the same float
  inputs; per-component sums and L2 norms in plain FP32, FP32 Neumaier and float-float against an FP64
  compensated reduction of the same float values, each within a stated bound -- the float-float bound is
  one a plain FP32 sum fails.
* The FFT batch of 4096 signals with lengths drawn from 2^6..2^13 and N(0,1) entries from Philox seed 7,
  in all four precisions, forward and inverse: 16 signals against the long-double DFT with the
  reference's bound (tests/test_Fft_common.h:79-83), the compensated complex64 modes held to the output
  rounding alone, and the round trip of every signal.
"""
import math

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu

WINDOW = 1 << 14


def windows(total):
    """Head, an odd middle offset, and the end of a stream of `total` elements."""
    return [0, total // 2 - 12345, total - WINDOW]


@pytest.mark.parametrize("kind,dtype,total,oracle", [
    (ae.RNG_BITS32, torch.int32, 1 << 30, "fill_bits32"),
    (ae.RNG_UNIFORM_F64, torch.float64, 1 << 29, "fill_uniform_f64"),
    (ae.RNG_NORMAL_F64, torch.float64, 1 << 28, "fill_normal_f64"),
], ids=["bits32_2^30", "uniform_f64_2^29", "normal_f64_2^28"])
def test_philox_seed42_full_streams(dev, orc, kind, dtype, total, oracle):
    out = torch.empty(total, dtype=dtype, device="cuda")
    ae.rng_fill(kind, 42, out)
    torch.cuda.synchronize()
    for first in windows(total):
        got = out[first:first + WINDOW].cpu().numpy().view(np.uint8)
        want = getattr(orc, oracle)(42, WINDOW, first).view(np.uint8)
        assert np.array_equal(got, want), (oracle, first)
    if kind == ae.RNG_UNIFORM_F64:  # the open interval (0, 1) (random/detail/Backend.h:121-134)
        assert out.min().item() > 0.0 and out.max().item() < 1.0
    if kind == ae.RNG_NORMAL_F64:  # 2^28 standard normals: mean and variance within 6 sigma
        m, v = out.mean().item(), out.var().item()
        assert abs(m) < 6 / math.sqrt(total) and abs(v - 1) < 6 * math.sqrt(2 / total)
    del out


def _uniform_arrays(n, dt, seeds=(42, 43, 44)):
    arrays = []
    for s in seeds:
        a = ae.Array(n, 3, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        arrays.append(a)
    return arrays


def _window_planes(arr, first):
    """(3, WINDOW) numpy copy of samples [first, first + WINDOW) of each component."""
    return np.ascontiguousarray(arr.data[:, first:first + WINDOW].cpu().numpy())


def test_fused_fp64_and_vec3_at_2p27(dev, orc):
    n = 1 << 27
    a, b, c = _uniform_arrays(n, torch.float64)
    d = ae.Array(n, 3, torch.float64)
    tree = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))
    sums = E.assign(d, tree, ae.SUM_NEUMAIER)
    e = torch.empty(n, dtype=torch.float64, device="cuda")
    ae.eval_vec3geo(a, b, c, e)
    torch.cuda.synchronize()
    for first in windows(n):
        A, B, C = (_window_planes(x, first).reshape(-1) for x in (a, b, c))
        want_d = orc.eval_cfg0(A, B, C, 3, WINDOW, WINDOW, mathmode=1).reshape(3, WINDOW)
        assert np.array_equal(_window_planes(d, first).view(np.uint64), want_d.view(np.uint64)), first
        want_e = orc.eval_vec3geo(A, B, C, WINDOW, WINDOW)
        assert np.array_equal(e[first:first + WINDOW].cpu().numpy().view(np.uint64), want_e.view(np.uint64)), first
    # the fused sums equal the standalone reduction of d (one size-only shape)
    assert torch.equal(sums, ae.reduce(d, ae.SUM_NEUMAIER))
    # Vec3 everywhere: the reference's right-nested folds as separate torch operations
    A, B, C = a.data[:, :n], b.data[:, :n], c.data[:, :n]
    x0 = A[1] * B[2] - A[2] * B[1]
    x1 = A[2] * B[0] - A[0] * B[2]
    x2 = A[0] * B[1] - A[1] * B[0]
    dxc = (x2 * C[2]) + ((x1 * C[1]) + (x0 * C[0]))
    daa = (A[2] * A[2]) + ((A[1] * A[1]) + (A[0] * A[0]))
    ref = dxc + torch.sqrt(daa + 1.0)
    ulps = ((e.view(torch.int64) - ref.view(torch.int64)).abs()).max().item()
    assert ulps <= 1, ulps


def test_fp32_at_2p28_error_against_fp64(dev):
    n = 1 << 28
    a, b, c = _uniform_arrays(n, torch.float32)
    d = ae.Array(n, 3, torch.float32)
    results = {}
    for name, mode in (("plain", ae.SUM_PLAIN), ("neumaier", ae.SUM_NEUMAIER), ("floatfloat", ae.SUM_FLOATFLOAT)):
        s = ae.eval_cfg0(a, b, c, d, mode).double().sum(1).cpu()
        q = ae.reduce(d, mode, square=True).double().sum(1).cpu()
        results[name] = (s, q)
    # FP64 references over the same float values of d, and the FP64 evaluation of the same float inputs
    d64 = ae.Array(n, 3, torch.float64)
    d64.data[:, :n].copy_(d.data[:, :n])
    ref_s = ae.reduce(d64, ae.SUM_NEUMAIER).sum(1).cpu()
    ref_q = ae.reduce(d64, ae.SUM_NEUMAIER, square=True).sum(1).cpu()
    abs_sum = d64.data[:, :n].abs().sum(1).cpu()
    del d64
    wide = []
    for x in (a, b, c):
        w = ae.Array(n, 3, torch.float64)
        w.data[:, :n].copy_(x.data[:, :n])
        wide.append(w)
    exact = ae.Array(n, 3, torch.float64)
    ae.eval_cfg0(*wide, exact, ae.SUM_NONE)
    # |d32 - d64| <= 4 ulp of the terms' scale |a| + |b c| + |sin a exp(-b^2)| <= |a| + |b c| + 1
    A, B, C = (w.data[:, :n] for w in wide)
    scale = A.abs() + (B * C).abs() + 1.0
    worst = ((d.data[:, :n].double() - exact.data[:, :n]).abs() / scale).max().item()
    assert worst <= 4 * 2.0**-24, worst
    del wide, exact, A, B, C, scale
    u = 2.0**-24
    for k in range(3):
        S, Q, T = ref_s[k].item(), ref_q[k].item(), abs_sum[k].item()
        err = {name: abs(results[name][0][k].item() - S) for name in results}
        norm_err = {name: abs(math.sqrt(results[name][1][k].item()) - math.sqrt(Q)) / math.sqrt(Q) for name in results}
        # plain FP32 with the pairwise tile tree: (2 + log2 n) u sum|d|
        assert err["plain"] <= (2 + 28) * u * T and norm_err["plain"] <= (2 + 28) * u
        # compensated FP32: 2u|S| plus second-order terms
        assert err["neumaier"] <= 2 * u * abs(S) + 64 * u * u * T
        # float-float: a 2^-44 relative carrier; plain FP32 (about 1e-2 absolute here) cannot meet this
        ff_bound = 2.0**-44 * T
        assert err["floatfloat"] <= ff_bound and norm_err["floatfloat"] <= 2.0**-40
        assert err["plain"] > ff_bound, "the float-float bound must separate compensated from plain sums"


PRECISIONS = [(0, "c128"), (1, "c64"), (2, "c64_floatfloat"), (3, "c64_wide")]


def _fft_batch_input():
    g = ae.Generator.from_seed(7)
    lengths = (1 << np.random.default_rng(7).integers(6, 14, 4096)).astype(np.uint64)
    total = int(lengths.sum())
    z = ae.Array(total, 2, torch.float64)
    E.assign(z, E.normal(g, 2, 0.0, 1.0))
    x = torch.complex(z.data[0, :total], z.data[1, :total]).contiguous()
    return lengths, x


@pytest.mark.parametrize("prec,name", PRECISIONS, ids=[p[1] for p in PRECISIONS])
def test_fft_batch_4096_seed7(dev, orc, prec, name):
    lengths, x128 = _fft_batch_input()
    offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
    dt = torch.complex128 if prec == 0 else torch.complex64
    eps = 2.0**-52 if prec == 0 else 2.0**-23
    x = x128.to(dt)
    fwd, inv = torch.empty_like(x), torch.empty_like(x)
    plan = ae.FftBatch(lengths, prec=prec)
    plan(torch.view_as_real(x), torch.view_as_real(fwd), True, 0)
    plan(torch.view_as_real(fwd), torch.view_as_real(inv), False, 0)
    torch.cuda.synchronize()
    plan.close()
    xs = x.to(torch.complex128).cpu().numpy()
    X, back = fwd.to(torch.complex128).cpu().numpy(), inv.to(torch.complex128).cpu().numpy()
    # 16 signals (every 256th) against the long-double DFT of the stored input
    for i in range(0, 4096, 256):
        nn, o = int(lengths[i]), int(offsets[i])
        seg = xs[o:o + nn]
        ref = orc.dft(seg, forward=True)
        bound = (8 + 4 * math.log2(nn + 1)) * eps * np.linalg.norm(seg)
        err = np.max(np.abs(X[o:o + nn] - ref))
        assert err <= bound, (i, nn, err, bound)
        if prec in (2, 3):  # compensated: only the rounding of the output to complex64 is left
            assert err <= 4 * 2.0**-24 * np.max(np.abs(ref)), (i, nn, err)
        ref_inv = orc.dft(X[o:o + nn], forward=False, scale=1.0 / nn)
        inv_bound = (8 + 4 * math.log2(nn + 1)) * eps * np.linalg.norm(X[o:o + nn]) / nn
        assert np.max(np.abs(back[o:o + nn] - ref_inv)) <= inv_bound + eps * np.max(np.abs(ref_inv))
    # every signal's round trip: 2 (8 + 4 log2 N) eps ||x||_2
    err = np.abs(back - xs)
    starts = offsets.astype(np.intp)
    worst = np.maximum.reduceat(err, starts)
    norms = np.sqrt(np.add.reduceat(np.abs(xs) ** 2, starts))
    bound = 2 * (8 + 4 * np.log2(lengths.astype(np.float64) + 1)) * eps * norms
    assert np.all(worst <= bound), int(np.argmax(worst / bound))
