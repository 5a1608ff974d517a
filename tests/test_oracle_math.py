"""Pins the oracle's own math (the algorithm the device kernels run).

Fixtures: tests/golden/bandmath_{sin,cos,exp,log}.npz -- the MPFR (256-bit,
correctly rounded) rows of the reference's tests/bandmath/golden_*.h.
Stated bound: <= 1 ulp of the correctly rounded result for double; the
float routines (double evaluation with float-length polynomials, one
rounding) within 0.51 float ulp of the exact mpmath value over their whole
domains, plus their special values.
"""
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import GOLDEN

ULP_BOUND_F64 = 1.0


def ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    sp = np.abs(np.spacing(ref))
    sp = np.where(sp == 0, 5e-324, sp)
    with np.errstate(all="ignore"):
        return np.abs(got - ref) / sp


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "log"])
def test_mpfr_golden_rows(orc, fn):
    z = np.load(os.path.join(GOLDEN, f"bandmath_{fn}.npz"))
    x = z["x_bits"].view(np.float64)
    ref = z["ref_bits"].view(np.float64)
    y = orc.math64(fn, x, mode=1)
    fin = np.isfinite(ref)
    assert np.max(ulp_err(y[fin], ref[fin])) <= ULP_BOUND_F64
    # non-finite references: NaN stays NaN, infinities and zeros exact
    nf = ~fin
    assert np.all((np.isnan(ref[nf]) & np.isnan(y[nf])) | (ref[nf] == y[nf]))


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_trig_specials(orc, fn):
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -1e-300, 1e300, -1e308, 2.0**1023], np.float64)
    y = orc.math64(fn, x, mode=1)
    if fn == "sin":
        assert y[0] == 0 and not np.signbit(y[0]) and y[1] == 0 and np.signbit(y[1])
        assert y[5] == 5e-324 and y[6] == -1e-300
    else:
        assert y[0] == 1 and y[1] == 1
    assert np.all(np.isnan(y[2:5]))
    lib = orc.math64(fn, x[7:], mode=0)
    assert np.max(ulp_err(y[7:], lib)) <= 1.0


def test_exp_log_specials(orc):
    x = np.array([0.0, -0.0, 1.0, np.inf, -np.inf, np.nan, 709.78, 710.0, -745.0, -746.0, 1e-310])
    e = orc.math64("exp", x, mode=1)
    assert e[0] == 1 and e[1] == 1 and e[3] == np.inf and e[4] == 0 and np.isnan(e[5])
    assert np.isfinite(e[6]) and e[7] == np.inf and e[8] > 0 and e[9] == 0
    l = orc.math64("log", np.array([0.0, -1.0, 1.0, np.inf, np.nan, 5e-324, 1e-310]), mode=1)
    assert l[0] == -np.inf and np.isnan(l[1]) and l[2] == 0 and l[3] == np.inf and np.isnan(l[4])
    assert l[5] == np.log(5e-324) and abs(l[6] - np.log(1e-310)) <= np.spacing(abs(np.log(1e-310)))


def _mp_ref(fn, xs):
    f = {"sin": mp.sin, "cos": mp.cos, "exp": mp.exp, "log": mp.log}[fn]
    with mp.workprec(200):
        return np.array([float(f(mp.mpf(float(v)))) for v in xs])


@pytest.mark.parametrize("fn,lo,hi", [("sin", -8.0, 8.0), ("cos", -1e6, 1e6), ("exp", -700.0, 700.0),
                                      ("log", 1e-300, 1e300)])
def test_random_vs_mpmath(orc, fn, lo, hi):
    rng = np.random.default_rng(11)
    if fn == "log":
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 2000))
    else:
        x = rng.uniform(lo, hi, 2000)
    y = orc.math64(fn, x, mode=1)
    assert np.max(ulp_err(y, _mp_ref(fn, x))) <= ULP_BOUND_F64


def test_payne_hanek_range(orc):
    # huge arguments take the Payne-Hanek path; compare with a 2000-bit mpmath reference
    rng = np.random.default_rng(3)
    x = np.ldexp(rng.uniform(1, 2, 300), rng.integers(21, 1023, 300))
    x[::2] *= -1
    y = orc.math64("sin", x, mode=1)
    with mp.workprec(2400):
        ref = np.array([float(mp.sin(mp.mpf(float(v)))) for v in x])
    assert np.max(ulp_err(y, ref)) <= ULP_BOUND_F64


FLOAT_DOMAINS = {
    # sin/cos up to float max (Payne-Hanek above 2^20); exp down into the subnormal results and up to
    # overflow; log over subnormal to max inputs
    "sin": lambda r, n: 10.0 ** r.uniform(-8, 38.5, n) * np.where(r.random(n) < 0.5, -1, 1),
    "cos": lambda r, n: 10.0 ** r.uniform(-8, 38.5, n) * np.where(r.random(n) < 0.5, -1, 1),
    "exp": lambda r, n: np.concatenate([r.uniform(-104, 88.7, n // 2),
                                        10.0 ** r.uniform(-9, 0, n - n // 2) * np.where(r.random(n - n // 2) < 0.5, -1, 1)]),
    "log": lambda r, n: np.concatenate([10.0 ** r.uniform(-45, 38.5, n // 2), 1.0 + r.uniform(-0.3, 0.42, n - n // 2)]),
}


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "log"])
def test_float_routines(orc, fn):
    """Float routines (argument in double, float-length polynomials, one rounding): within 0.51 ulp of the
    exact value (mpmath) over their whole domains -- 2^-30 relative before the rounding."""
    rng = np.random.default_rng(5)
    x = FLOAT_DOMAINS[fn](rng, 4000).astype(np.float32)
    x = x[np.isfinite(x) & (x != 0)]
    y = orc.math32(fn, x, mode=1)
    ref = _mp_ref(fn, x.astype(np.float64))
    ok = np.isfinite(ref) & (np.abs(ref) <= np.finfo(np.float32).max)
    r32 = ref[ok].astype(np.float32)
    sp = np.abs(np.spacing(r32)).astype(np.float64)
    err = np.abs(y[ok].astype(np.float64) - ref[ok]) / sp
    assert np.max(err) <= 0.51
    assert np.mean(y[ok] == r32) >= 0.999  # correctly rounded but for near-ties


def test_float_specials(orc):
    f32 = lambda v: np.array(v, np.float32)  # noqa: E731
    s = orc.math32("sin", f32([0.0, -0.0, 1e-40, -1e-40, np.inf, -np.inf, np.nan, 3.4e38]))
    assert s[0] == 0 and not np.signbit(s[0]) and s[1] == 0 and np.signbit(s[1])
    assert s[2] == np.float32(1e-40) and s[3] == np.float32(-1e-40) and np.all(np.isnan(s[4:7]))
    assert abs(s[7]) <= 1
    c = orc.math32("cos", f32([0.0, -0.0, np.inf, np.nan]))
    assert c[0] == 1 and c[1] == 1 and np.all(np.isnan(c[2:]))
    e = orc.math32("exp", f32([-np.inf, np.inf, np.nan, 0.0, 88.8, -103.5, -104.5, -87.5]))
    assert e[0] == 0 and e[1] == np.inf and np.isnan(e[2]) and e[3] == 1 and e[4] == np.inf
    assert e[5] == np.float32(np.exp(np.float64(np.float32(-103.5)))) and e[5] > 0  # subnormal, rounded once
    assert e[6] == 0 and e[7] == np.float32(np.exp(np.float64(np.float32(-87.5))))
    l = orc.math32("log", f32([0.0, -0.0, -1.0, np.inf, np.nan, 1.0, 1e-45]))
    assert l[0] == -np.inf and l[1] == -np.inf and np.isnan(l[2]) and l[3] == np.inf and np.isnan(l[4])
    assert l[5] == 0 and l[6] == np.float32(np.log(np.float64(np.float32(1e-45))))
