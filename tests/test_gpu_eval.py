"""GPU parity: fused evaluation, Vec3 geometry, reductions and math vs the oracle.

Bit for bit against the oracle's restatement (own-math route for the
transcendentals; the fixed reduction shape for the sums). At full size
(2^27 samples) only size-independent properties are checked.
"""
import math

import numpy as np
import pytest
import torch

import aether_amd as ae

pytestmark = pytest.mark.gpu


def filled(n, dt, seeds=(42, 43, 44), ncomp=3):
    arrs = []
    for s in seeds:
        a = ae.Array(n, ncomp, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        arrs.append(a)
    return arrs


def np_dt(dt):
    return np.float64 if dt == torch.float64 else np.float32


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [0, 1, 31, 2000, 2048, 2049, 100_003, 1 << 20])
@pytest.mark.parametrize("mode", [ae.SUM_NEUMAIER, ae.SUM_PLAIN, ae.SUM_FLOATFLOAT, ae.SUM_NONE])
def test_cfg0_parity(dev, orc, dt, n, mode):
    if mode == ae.SUM_FLOATFLOAT and dt == torch.float64:
        with pytest.raises(ae.AetherError) as e:
            a, b, c = filled(8, dt)
            ae.eval_cfg0(a, b, c, ae.Array(8, 3, dt), mode)
        assert e.value.code == 2
        return
    a, b, c = filled(n, dt)
    d = ae.Array(n, 3, dt)
    sums = ae.eval_cfg0(a, b, c, d, mode)
    torch.cuda.synchronize()
    pitch = a.capacity
    A, B, C = (x.data.cpu().numpy().reshape(-1) for x in (a, b, c))
    exp = orc.eval_cfg0(A, B, C, 3, n, pitch, mathmode=1)
    got = d.data.cpu().numpy().reshape(-1)
    if n:
        g3, e3 = got.reshape(3, -1)[:, :n], exp.reshape(3, -1)[:, :n]
        assert np.array_equal(g3.view(np.uint8), e3.view(np.uint8))
    if mode == ae.SUM_NONE:
        return
    s = sums.cpu().numpy()
    om = {ae.SUM_NEUMAIER: "neumaier", ae.SUM_PLAIN: "plain", ae.SUM_FLOATFLOAT: "ff"}[mode]
    for k in range(3):
        plane = np.ascontiguousarray(exp.reshape(3, -1)[k, :n])
        h, cc = orc.device_sum(plane, mode=om)
        assert s[k, 0] == h and s[k, 1] == cc, (k, s[k], h, cc)


def test_cfg0_unaligned_views(dev, orc):
    # a plane pointer off the 16-byte grid takes the scalar path with the same bits
    n = 5000
    a, b, c = filled(n, torch.float64)
    d = ae.Array(n, 3, torch.float64)
    raw = torch.empty(3 * a.capacity + 1, dtype=torch.float64, device="cuda")
    dv = raw[1:].view(3, a.capacity)
    sums = torch.empty((3, 2), dtype=torch.float64, device="cuda")
    ae._lib.call("aeth_eval_cfg0", ae.F64, a.ptr(), b.ptr(), c.ptr(), dv.data_ptr(), 3, n, a.capacity,
                 ae.SUM_NEUMAIER, sums.data_ptr(), ae.stream_handle())
    ref_sums = ae.eval_cfg0(a, b, c, d, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert torch.equal(dv[:, :n].cpu(), d.data[:, :n].cpu())
    assert torch.equal(sums.cpu(), ref_sums.cpu())


def test_cfg0_full_size_properties(dev):
    # 2^27 samples: sums agree with an FP64 re-sum of d, and d is deterministic across runs
    n = 1 << 27
    a, b, c = filled(n, torch.float64)
    d = ae.Array(n, 3, torch.float64)
    s1 = ae.eval_cfg0(a, b, c, d, ae.SUM_NEUMAIER).clone()
    d1 = d.data[:, :4096].clone()
    s2 = ae.eval_cfg0(a, b, c, d, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert torch.equal(s1, s2) and torch.equal(d1, d.data[:, :4096])
    # the fused sum equals the standalone reduction of d (same size-only shape), bit for bit
    s3 = ae.reduce(d, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert torch.equal(s1, s3)
    for k in range(3):
        plane = d.data[k, :n]
        ref = torch.sum(plane).item()
        got = s1[k, 0].item() + s1[k, 1].item()
        # torch's pairwise FP64 sum is the less accurate side: log2(n) eps sum|x|
        assert abs(got - ref) <= 30 * 2.0**-52 * plane.abs().sum().item()
    # spot check against the closed form on a few samples
    i = torch.tensor([0, 12345, n // 2, n - 1], device="cuda")
    A, B, C = a.data[0, i].cpu(), b.data[0, i].cpu(), c.data[0, i].cpu()
    ref = A + B * C + torch.sin(A) * torch.exp(-(B * B))
    assert torch.allclose(d.data[0, i].cpu(), ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 999, 1024, 1025, 2000, 65537, 8_000_001])
def test_vec3geo_parity(dev, orc, n):
    a, b, c = filled(n, torch.float64)
    e = torch.empty(n, dtype=torch.float64, device="cuda")
    ae.eval_vec3geo(a, b, c, e)
    torch.cuda.synchronize()
    A, B, C = (x.data.cpu().numpy().reshape(-1) for x in (a, b, c))
    exp = orc.eval_vec3geo(A, B, C, n, a.capacity)
    assert np.array_equal(e.cpu().numpy().view(np.uint64), exp.view(np.uint64))


def test_vec3geo_unaligned_output(dev, orc):
    # an output one element off 16-byte alignment takes the element-wise path
    n = 5001
    a, b, c = filled(n, torch.float64)
    buf = torch.empty(n + 1, dtype=torch.float64, device="cuda")
    ae.eval_vec3geo(a, b, c, buf[1:])
    torch.cuda.synchronize()
    A, B, C = (x.data.cpu().numpy().reshape(-1) for x in (a, b, c))
    exp = orc.eval_vec3geo(A, B, C, n, a.capacity)
    assert np.array_equal(buf[1:].cpu().numpy().view(np.uint64), exp.view(np.uint64))


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("square", [False, True])
@pytest.mark.parametrize("mode", [ae.SUM_NEUMAIER, ae.SUM_PLAIN, ae.SUM_FLOATFLOAT])
@pytest.mark.parametrize("n", [0, 5, 2048, 300_001])
def test_reduce_parity(dev, orc, dt, square, mode, n):
    if mode == ae.SUM_FLOATFLOAT and dt == torch.float64:
        # float-float accumulation is an FP32 mode: FP64 data are refused as unsupported (code 2)
        (x,) = filled(max(n, 1), dt, seeds=(9,))
        with pytest.raises(ae.AetherError) as e:
            ae.reduce(x, mode, square)
        assert e.value.code == 2
        return
    (x,) = filled(n, dt, seeds=(9,))
    out = ae.reduce(x, mode, square)
    torch.cuda.synchronize()
    om = {ae.SUM_NEUMAIER: "neumaier", ae.SUM_PLAIN: "plain", ae.SUM_FLOATFLOAT: "ff"}[mode]
    X = x.data.cpu().numpy()
    for k in range(3):
        h, c = orc.device_sum(np.ascontiguousarray(X[k, :n]), square=square, mode=om)
        assert out[k, 0].item() == h and out[k, 1].item() == c


@pytest.mark.parametrize("with_nonfinite", [False, True])
def test_reduce_floatfloat_special_values(dev, orc, with_nonfinite):
    """Float-float sums over planes thick with signed zeros, subnormals and values of every magnitude (and,
    in the second case, infinities and NaNs): the (hi, lo) pair of every plane is the oracle's bit for bit,
    zero signs and NaN included, plain and squared."""
    rng = np.random.default_rng(21)
    n = 70_001
    x = ae.Array(n, 3, torch.float32)
    vals = (rng.standard_normal((3, n)) * 10.0 ** rng.integers(-44, 38, (3, n))).astype(np.float32)
    pick = rng.integers(0, 8, (3, n))
    vals[pick == 0] = -0.0
    vals[pick == 1] = 0.0
    vals[pick == 2] = np.float32(1e-45) * rng.integers(-50, 50, int((pick == 2).sum())).astype(np.float32)
    if with_nonfinite:
        for k, v in enumerate([np.inf, -np.inf, np.nan]):
            vals[k, rng.integers(0, n, 3)] = v
    x.data[:, :n] = torch.from_numpy(vals).to("cuda")
    for square in (False, True):
        out = ae.reduce(x, ae.SUM_FLOATFLOAT, square).cpu().numpy()
        for k in range(3):
            h, c = orc.device_sum(np.ascontiguousarray(vals[k]), square=square, mode="ff")
            got = out[k].astype(np.float32)
            want = np.array([h, c], dtype=np.float32)
            same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
            assert same.all(), (k, square, got, want)


def test_reduce_nonfinite(dev):
    x = ae.Array(5000, 1, torch.float64)
    x.data.fill_(1.0)
    x.data[0, 17] = float("inf")
    out = ae.reduce(x, ae.SUM_NEUMAIER).cpu()
    assert out[0, 0].item() + out[0, 1].item() == math.inf
    x.data[0, 400] = float("-inf")
    out = ae.reduce(x, ae.SUM_NEUMAIER).cpu()
    assert math.isnan(out[0, 0].item() + out[0, 1].item())


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "log"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_math_bit_exact(dev, orc, fn, dt):
    rng = np.random.default_rng(1)
    n = 200_000
    mag = 10.0 ** rng.uniform(-6, 5, n)
    x = mag * np.where(rng.random(n) < 0.5, -1, 1)
    if fn == "log":
        x = np.abs(x)
    if fn == "exp":
        x = np.clip(x, -700, 700) if dt == torch.float64 else np.clip(x, -104, 89)  # float: subnormal results too
    specials = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-300, 1e300, 2.0**60, -2.0**40, 5e-324, 1e-40, -1e-40,
                3e38, -103.5, 88.8]
    x = np.concatenate([x, specials]).astype(np_dt(dt))
    y = ae.math(fn, torch.from_numpy(x).to("cuda"))
    torch.cuda.synchronize()
    got = y.cpu().numpy()
    exp = orc.math64(fn, x) if dt == torch.float64 else orc.math32(fn, x)
    same = (got.view(np.uint8).reshape(len(x), -1) == exp.view(np.uint8).reshape(len(x), -1)).all(1)
    both_nan = np.isnan(got) & np.isnan(exp)
    assert np.all(same | both_nan)


def test_math_all4(dev, orc):
    rng = np.random.default_rng(2)
    x = (10.0 ** rng.uniform(-6, 5, 4096) * np.where(rng.random(4096) < 0.5, -1, 1)).astype(np.float32)
    y = ae.math("all4", torch.from_numpy(x).to("cuda")).cpu().numpy().reshape(4, -1)
    for k, fn in enumerate(["sin", "cos", "exp", "log"]):
        e = orc.math32(fn, x)
        assert np.all((y[k].view(np.uint32) == e.view(np.uint32)) | (np.isnan(y[k]) & np.isnan(e)))
