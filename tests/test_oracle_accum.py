"""Pins the oracle's compensated accumulation.

Mirrors the reference's tests/test_CompensatedSum_common.h:29-196 (exact-by-
construction fixtures, <= 1 ulp; plain sum fails; pairCombine_ bitwise
symmetric; residual kept; Inf/NaN rules) and checks the device reduction
shape against math.fsum (correctly rounded sum) on random data.
"""
import math

import numpy as np
import pytest


def make_fixture(seed, dtype):
    # +-x pairs with full mantissas plus residuals on the 2^-10 grid: exact sum known
    rng = np.random.default_rng(seed)
    xs = rng.uniform(1.0, 3.0, 64).astype(dtype)
    g = rng.integers(1, 5, 64)
    terms = np.concatenate([xs, -xs, (g * 0.0009765625).astype(dtype)])
    rng.shuffle(terms)
    return terms, float(g.sum()) * 0.0009765625


def ulp_err(got, exact, dtype):
    e = dtype(exact)
    ulp = float(np.nextafter(e, dtype(np.inf)) - e)
    return abs(float(got) - exact) / ulp


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tree_and_device_shape_within_one_ulp(orc, dtype):
    for seed in range(1, 9):
        terms, exact = make_fixture(seed, dtype)
        assert ulp_err(orc.tree_sum32(terms), exact, dtype) <= 1.0
        h, c = orc.device_sum(terms)
        assert ulp_err(dtype(h) + dtype(c), exact, dtype) <= 1.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plain_sum_fails_fixture(dtype):
    fails = 0
    for seed in range(1, 9):
        terms, exact = make_fixture(seed, dtype)
        s = dtype(0)
        for t in terms:
            s = dtype(s + t)
        fails += ulp_err(s, exact, dtype) > 1.0
    assert fails >= 6


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pair_combine_symmetric_bitwise(orc, dtype):
    rng = np.random.default_rng(7)
    for i in range(3000):
        ah = dtype(rng.uniform(-1, 1) * 2.0 ** (i % 20 - 10))
        bh = dtype(rng.uniform(-1, 1) * 2.0 ** ((i // 7) % 20 - 10))
        ac, bc = dtype(rng.uniform(-1, 1) * 1e-6), dtype(rng.uniform(-1, 1) * 1e-6)
        x = orc.pair_combine(ah, ac, bh, bc, dtype)
        y = orc.pair_combine(bh, bc, ah, ac, dtype)
        assert np.array(x, dtype).tobytes() == np.array(y, dtype).tobytes()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_residual_kept_and_nonfinite_rules(orc, dtype):
    big = float(np.ldexp(dtype(1), np.finfo(dtype).nmant + 4))
    oh, oc = orc.pair_combine(big, 0.0, 0.0009765625, 0.0, dtype)
    assert oh == big and oc == 0.0009765625
    h, c = 0.0, 0.0
    for t in (1.5, np.inf, 2.25):
        h, c = orc.neumaier_add(h, c, t, dtype)
    assert h + c == np.inf
    h, c = 0.0, 0.0
    for t in (np.inf, 1.0, -np.inf):
        h, c = orc.neumaier_add(h, c, t, dtype)
    assert math.isnan(h + c)
    h, c = 0.0, 0.0
    for t in (1.0, np.nan, 1.0):
        h, c = orc.neumaier_add(h, c, t, dtype)
    assert math.isnan(h + c)


@pytest.mark.parametrize("n", [0, 1, 31, 2047, 2048, 2049, 100_000, 3_000_001])
def test_device_shape_vs_fsum_f64(orc, n):
    rng = np.random.default_rng(n + 1)
    x = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-3, 3, n)
    h, c = orc.device_sum(x)
    exact = math.fsum(x)
    assert abs((h + c) - exact) <= 2 * np.spacing(abs(exact)) + 1e-300
    # sum of squares
    h2, c2 = orc.device_sum(x, square=True)
    exact2 = math.fsum((x * x).tolist())
    assert abs((h2 + c2) - exact2) <= 4 * np.spacing(exact2) + 1e-300


@pytest.mark.parametrize("n", [1000, 2048 * 33 + 5, 1 << 20])
def test_float_modes_error_vs_fp64(orc, n):
    # plain FP32 is worst; Neumaier and float-float land within a few float ulps of the exact sum
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    exact = math.fsum(x.astype(np.float64))
    ph, _ = orc.device_sum(x, mode="plain")
    nh, nc = orc.device_sum(x, mode="neumaier")
    fh, fl = orc.device_sum(x, mode="ff")
    ff = float(fh) + float(fl)
    assert abs(ff - exact) <= 1e-9 * max(1.0, abs(exact)) + 1e-6
    assert abs(float(nh) + float(nc) - exact) <= 2 * abs(float(np.spacing(np.float32(exact)))) + 1e-6
    # plain FP32 accumulates n * eps-scale error; it is never better than the float-float result
    assert abs(float(ph) - exact) >= abs(ff - exact) - 1e-12


def test_device_shape_depends_on_size_only(orc):
    # the same values give the same bits whatever surrounds them (no launch dependence)
    x = np.random.default_rng(1).uniform(-1, 1, 50_000)
    assert orc.device_sum(x) == orc.device_sum(x.copy())


def test_block_sum_matches_fsum_and_the_serial_walk(orc):
    """The host's parallel compensated sum (per-thread Neumaier blocks combined in block order): one
    block is the serial Neumaier walk exactly; any block count stays within 2 ulp of math.fsum."""
    import math

    rng = np.random.default_rng(21)
    x = rng.standard_normal(300_001) * np.exp(rng.uniform(-20, 20, 300_001))
    h1, c1 = orc.block_sum(x, 1)
    h, c = 0.0, 0.0
    for t in x[:5000]:
        h, c = orc.neumaier_add(h, c, t)
    hs, cs = orc.block_sum(x[:5000], 1)
    assert (hs, cs) == (h, c)
    exact = math.fsum(x)
    for nt in (1, 2, 3, 8):
        hh, cc = orc.block_sum(x, nt)
        assert abs((hh + cc) - exact) <= 2 * math.ulp(exact)
