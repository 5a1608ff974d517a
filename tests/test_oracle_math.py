"""Pins the oracle's own math (the algorithm the device kernels run).

Fixtures: tests/golden/bandmath_{sin,cos,exp,log}.npz -- the MPFR (256-bit,
correctly rounded) rows of the reference's tests/bandmath/golden_*.h.
Stated bound: <= 1 ulp of the correctly rounded result for double; the
float routines (double evaluation, one rounding) are checked against the
correctly rounded float of an mpmath value on random samples.
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


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "log"])
def test_float_routines(orc, fn):
    rng = np.random.default_rng(5)
    mag = 10.0 ** rng.uniform(-6, 5, 3000)
    x = (mag * np.where(rng.random(3000) < 0.5, -1, 1)).astype(np.float32)
    if fn == "log":
        x = np.abs(x)
    if fn == "exp":
        x = np.clip(x, -87, 88).astype(np.float32)
    y = orc.math32(fn, x, mode=1)
    ref = _mp_ref(fn, x.astype(np.float64)).astype(np.float32)
    err = np.abs(y.astype(np.float64) - ref.astype(np.float64)) / np.abs(np.spacing(ref)).astype(np.float64)
    assert np.max(err) <= 1.0
