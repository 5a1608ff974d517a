"""Pins the oracle's own math (the algorithm the device kernels run).

Fixtures: tests/golden/bandmath_{sin,cos,exp,log}.npz -- the MPFR (256-bit,
correctly rounded) rows of the reference's tests/bandmath/golden_*.h.
Stated bound: <= 1 ulp of the correctly rounded result for double; the
float sin/cos/exp/log (float arithmetic with float-float steps, one final
rounding) within 0.8 float ulp of the exact value over their whole domains,
plus their special values.
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
    """Float routines (float arithmetic, float-float reduction and leading terms): within 0.8 ulp of the exact
    value (mpmath) over their whole domains, nearly all correctly rounded."""
    rng = np.random.default_rng(5)
    x = FLOAT_DOMAINS[fn](rng, 4000).astype(np.float32)
    x = x[np.isfinite(x) & (x != 0)]
    y = orc.math32(fn, x, mode=1)
    ref = _mp_ref(fn, x.astype(np.float64))
    ok = np.isfinite(ref) & (np.abs(ref) <= np.finfo(np.float32).max)
    r32 = ref[ok].astype(np.float32)
    sp = np.abs(np.spacing(r32)).astype(np.float64)
    err = np.abs(y[ok].astype(np.float64) - ref[ok]) / sp
    assert np.max(err) <= 0.8
    assert np.mean(y[ok] == r32) >= 0.98


@pytest.mark.parametrize("fn", ["sin", "cos", "exp", "log"])
def test_float_routines_dense(orc, fn):
    """A million arguments per routine against double libm (its error, 2^-52 relative, is 2^-29 of a float
    ulp): the same 0.8 ulp bound, and the double-path boundary of sin/cos (2^17) crossed on both sides."""
    rng = np.random.default_rng(17)
    n = 1 << 20
    dom = {"sin": (-7, 5.5), "cos": (-7, 5.5), "exp": (-2, 2), "log": (-40, 38)}[fn]
    x = 10.0 ** rng.uniform(*dom, n) * np.where(rng.random(n) < 0.5, -1, 1)
    if fn == "exp":
        x = np.clip(x, -103.9, 88.7)
    if fn == "log":
        x = np.abs(x)
    x = x.astype(np.float32)
    y = orc.math32(fn, x, mode=1).astype(np.float64)
    ref = {"sin": np.sin, "cos": np.cos, "exp": np.exp, "log": np.log}[fn](x.astype(np.float64))
    sp = np.abs(np.spacing(ref.astype(np.float32))).astype(np.float64)
    tiny = np.float64(2.0 ** -149)
    err = np.abs(y - ref) / np.maximum(sp, tiny)
    assert np.max(err) <= 0.8, (np.max(err), x[np.argmax(err)])
    assert np.mean(y == ref.astype(np.float32)) >= 0.98


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


# ---- the rest of the elementwise set: tan, atan, asin, acos, tanh, rsqrt, pow, sign, fmin, fmax, the ordered max
EXT = ["tan", "tanh", "atan", "asin", "acos", "rsqrt"]


@pytest.mark.parametrize("fn", EXT)
def test_mpfr_golden_rows_extended(orc, fn):
    """<= 1 ulp of the correctly rounded value on every row of the reference's MPFR corpus (in-domain,
    degraded and rejected rows alike); non-finite references exactly."""
    z = np.load(os.path.join(GOLDEN, f"bandmath_{fn}.npz"))
    x = z["x_bits"].view(np.float64)
    ref = z["ref_bits"].view(np.float64)
    y = orc.math64(fn, x)
    fin = np.isfinite(ref)
    assert np.max(ulp_err(y[fin], ref[fin])) <= ULP_BOUND_F64
    nf = ~fin
    assert np.all((np.isnan(ref[nf]) & np.isnan(y[nf])) | (ref[nf] == y[nf]))


def test_pow_mpfr_golden_rows(orc):
    z = np.load(os.path.join(GOLDEN, "bandmath_pow.npz"))
    x, yy, ref = z["x_bits"].view(np.float64), z["y_bits"].view(np.float64), z["ref_bits"].view(np.float64)
    got = orc.math2_64("pow", x, yy)
    fin = np.isfinite(ref)
    assert np.max(ulp_err(got[fin], ref[fin])) <= ULP_BOUND_F64
    nf = ~fin
    assert np.all((np.isnan(ref[nf]) & np.isnan(got[nf])) | (ref[nf] == got[nf]))


def _cr(f, *args):
    mp.mp.prec = 200
    return float(f(*[mp.mpf(float(a)) for a in args]))


@pytest.mark.parametrize("fn,mf,xs", [
    ("tan", mp.tan, lambda r: np.concatenate([r.uniform(-2, 2, 300), 10 ** r.uniform(-9, 7, 300) * r.choice([-1, 1], 300)])),
    ("atan", mp.atan, lambda r: np.concatenate([r.uniform(-2, 2, 300), 10 ** r.uniform(-12, 12, 300) * r.choice([-1, 1], 300)])),
    ("asin", mp.asin, lambda r: np.concatenate([r.uniform(-1, 1, 300), 1 - 10 ** r.uniform(-16, -1, 150),
                                                 -1 + 10 ** r.uniform(-16, -1, 150)])),
    ("acos", mp.acos, lambda r: np.concatenate([r.uniform(-1, 1, 300), 1 - 10 ** r.uniform(-16, -1, 150),
                                                 -1 + 10 ** r.uniform(-16, -1, 150)])),
    ("tanh", mp.tanh, lambda r: np.concatenate([r.uniform(-3, 3, 300), 10 ** r.uniform(-12, 1.4, 300) * r.choice([-1, 1], 300)])),
])
def test_extended_random_vs_mpmath(orc, fn, mf, xs):
    """Seeded random inputs over each function's whole range: within 1 ulp of the correctly rounded value
    (mpmath at 200 bits)."""
    x = xs(np.random.default_rng(11))
    y = orc.math64(fn, x)
    ref = np.array([_cr(mf, v) for v in x])
    assert np.max(ulp_err(y, ref)) <= 1.0


def test_pow_random_vs_mpmath(orc):
    """Bases over the whole exponent range, moderate and huge exponents, bases next to 1 with exponents up
    to 1e8 (where y log x must carry ~2^-63 relative accuracy): within 1 ulp wherever the result is a
    normal double."""
    r = np.random.default_rng(12)
    x = np.concatenate([10 ** r.uniform(-300, 300, 200), r.uniform(0, 4, 200), 1 + r.uniform(-1e-6, 1e-6, 200)])
    y = np.concatenate([r.uniform(-1, 1, 200), r.uniform(-200, 200, 200), r.uniform(-1e8, 1e8, 200)])
    got = orc.math2_64("pow", x, y)
    mp.mp.prec = 200
    for a, b, g in zip(x, y, got):
        ref = mp.power(mp.mpf(float(a)), mp.mpf(float(b)))
        if not (mp.mpf(2) ** -1022 < abs(ref) < mp.mpf(2) ** 1024):
            continue
        assert ulp_err(np.array([g]), np.array([float(ref)]))[0] <= 1.0, (a, b)


def test_pow_special_cases_follow_c99(orc):
    """C99 Annex F.9.4.4 (what the reference's std::pow and the device pow give) against numpy's power on
    every pairing of special bases and exponents."""
    vals = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0, np.inf, -np.inf, np.nan, 1e-310, -1e-310,
                     2.5, -2.5, 1e308, -1e308, 2.0 ** 53, -(2.0 ** 53) - 2, 2.0 ** 53 + 2])
    x, y = np.meshgrid(vals, vals)
    x, y = x.ravel(), y.ravel()
    got = orc.math2_64("pow", x, y)
    with np.errstate(all="ignore"):
        want = np.power(x, y)
    special = np.isin(x, [0.0, 1.0, -1.0, np.inf, -np.inf]) | np.isin(y, [0.0, 1.0, -1.0, np.inf, -np.inf])
    special |= np.isnan(x) | np.isnan(y) | ~np.isfinite(want) | (want == 0)
    same = (np.isnan(got) & np.isnan(want)) | ((got == want) & (np.signbit(got) == np.signbit(want)))
    assert np.all(same[special]), list(zip(x[~same], y[~same], got[~same], want[~same]))[:8]
    # the rest (finite, non-trivial) are ordinary values: within 1 ulp, with the sign numpy gives
    rest = ~special
    assert np.all(np.signbit(got[rest]) == np.signbit(want[rest]))
    assert np.max(ulp_err(got[rest], want[rest])) <= 1.0


def test_extended_specials(orc):
    inf, nan = np.inf, np.nan
    assert np.array_equal(orc.math64("tan", np.array([0.0, -0.0])), [0.0, -0.0])
    assert np.signbit(orc.math64("tan", np.array([-0.0]))[0])
    assert np.all(np.isnan(orc.math64("tan", np.array([inf, -inf, nan]))))
    at = orc.math64("atan", np.array([inf, -inf, 0.0, -0.0, 1.0, nan, 1e-300]))
    assert at[0] == np.pi / 2 and at[1] == -np.pi / 2 and at[2] == 0 and np.signbit(at[3])
    assert at[4] == np.pi / 4 and np.isnan(at[5]) and at[6] == 1e-300
    for fn in ("asin", "acos"):
        assert np.all(np.isnan(orc.math64(fn, np.array([1.0000000000000002, -1.5, inf, nan]))))
    assert np.array_equal(orc.math64("asin", np.array([1.0, -1.0, 0.0])), [np.pi / 2, -np.pi / 2, 0.0])
    assert np.array_equal(orc.math64("acos", np.array([1.0, -1.0, 0.0])), [0.0, np.pi, np.pi / 2])
    th = orc.math64("tanh", np.array([inf, -inf, 30.0, -0.0, nan]))
    assert th[0] == 1 and th[1] == -1 and th[2] == 1 and np.signbit(th[3]) and np.isnan(th[4])
    rs = orc.math64("rsqrt", np.array([0.0, inf, -1.0, 4.0]))
    assert rs[0] == inf and rs[1] == 0 and np.isnan(rs[2]) and rs[3] == 0.5
    sg = orc.math64("sign", np.array([3.0, -2.0, 0.0, -0.0, nan, -inf]))
    assert np.array_equal(sg, [1.0, -1.0, 0.0, 0.0, 0.0, -1.0]) and not np.signbit(sg[3])
    a = np.array([1.0, nan, 2.0, -0.0, 0.0, nan])
    b = np.array([2.0, 1.0, nan, 0.0, -0.0, nan])
    mn, mx, sg2 = orc.math2_64("fmin", a, b), orc.math2_64("fmax", a, b), orc.math2_64("selgt", a, b)
    assert mn[0] == 1 and mn[1] == 1 and mn[2] == 2 and not np.signbit(mn[3]) and np.signbit(mn[4]) and np.isnan(mn[5])
    assert mx[0] == 2 and mx[1] == 1 and mx[2] == 2 and not np.signbit(mx[3]) and np.signbit(mx[4]) and np.isnan(mx[5])
    assert sg2[0] == 2 and sg2[1] == 1 and np.isnan(sg2[2]) and not np.signbit(sg2[3]) and np.signbit(sg2[4])


FORM_BOUND = {"tan": 0.8, "atan": 0.51, "asin": 0.51, "acos": 0.51, "tanh": 0.65}


@pytest.mark.parametrize("fn", sorted(FORM_BOUND))
def test_float_forms_in_float_arithmetic(orc, fn):
    """tan/atan/asin/acos/tanh in float arithmetic: 2^20 arguments against double libm (its error is 2^-29
    of a float ulp) within the stated bound, over magnitudes from 10^-7 to the domain's edge (tan: to 10^6,
    across the double-reduction boundary 2^17), and 4000 of them against mpmath."""
    rng = np.random.default_rng(23)
    n = 1 << 20
    hi = {"tan": 6.0, "atan": 30.0, "asin": 0.0, "acos": 0.0, "tanh": 1.2}[fn]
    x = 10.0 ** rng.uniform(-7, hi, n) * np.where(rng.random(n) < 0.5, -1, 1)
    x = x.astype(np.float32)
    if fn in ("asin", "acos"):
        x = np.concatenate([x, (1 - 10.0 ** rng.uniform(-7.3, -0.2, n // 4)).astype(np.float32)])  # near 1
        x = np.concatenate([x, -x])
    y = orc.math32(fn, x, mode=1).astype(np.float64)
    ref = getattr(np, {"atan": "arctan", "asin": "arcsin", "acos": "arccos"}.get(fn, fn))(x.astype(np.float64))
    sp = np.abs(np.spacing(ref.astype(np.float32))).astype(np.float64)
    err = np.abs(y - ref) / np.maximum(sp, 2.0 ** -149)
    assert np.max(err) <= FORM_BOUND[fn], (np.max(err), x[np.argmax(err)])
    xs = x[:: len(x) // 4000][:4000]
    ys = orc.math32(fn, xs, mode=1).astype(np.float64)
    mf = {"tan": mp.tan, "atan": mp.atan, "asin": mp.asin, "acos": mp.acos, "tanh": mp.tanh}[fn]
    with mp.workdps(40):
        want = np.array([float(mf(mp.mpf(float(v)))) for v in xs])
    spm = np.abs(np.spacing(want.astype(np.float32))).astype(np.float64)
    assert np.max(np.abs(ys - want) / np.maximum(spm, 2.0 ** -149)) <= FORM_BOUND[fn]


def test_float_form_specials(orc):
    f32 = lambda v: np.array(v, np.float32)  # noqa: E731
    sp = f32([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 1.5, 1e-40, -1e-40])
    t = orc.math32("tan", sp)
    assert t[0] == 0 and np.signbit(t[1]) and np.isnan(t[2]) and np.isnan(t[3]) and np.isnan(t[4])
    assert t[8] == sp[8] and t[9] == sp[9]
    a = orc.math32("atan", sp)
    assert a[0] == 0 and np.signbit(a[1]) and a[2] == np.float32(np.pi / 2) and a[3] == -np.float32(np.pi / 2)
    assert np.isnan(a[4]) and a[5] == np.float32(np.pi / 4) and a[8] == sp[8]
    s = orc.math32("asin", sp)
    assert s[0] == 0 and np.signbit(s[1]) and np.isnan(s[2]) and np.isnan(s[7]) and s[5] == np.float32(np.pi / 2)
    assert s[6] == -np.float32(np.pi / 2) and s[8] == sp[8]
    c = orc.math32("acos", sp)
    assert c[5] == 0 and c[6] == np.float32(np.pi) and c[0] == np.float32(np.pi / 2) and np.isnan(c[7])
    h = orc.math32("tanh", sp)
    assert h[0] == 0 and np.signbit(h[1]) and h[2] == 1 and h[3] == -1 and np.isnan(h[4]) and h[8] == sp[8]
    assert orc.math32("tanh", f32([9.02, 20.0, 88.0]))[0] == 1


def test_float_pow_in_float_arithmetic(orc):
    """pow's float form: 3*2^18 pairs against double pow (its error is 2^-29 of a float ulp) -- x over the
    whole float range with y keeping |y ln x| <= 100, x within 10^-4 of 1 with |y| up to 10^6, x in [0, 4)
    with y in [-10, 10] -- within 0.8 ulp (subnormal results included); 2000 of them against mpmath."""
    rng = np.random.default_rng(29)
    n = 1 << 18
    x1 = 2.0 ** rng.uniform(-149, 128, n)
    y1 = rng.uniform(-1, 1, n) * 100.0 / np.maximum(np.abs(np.log(x1)), 1e-3)
    x2 = 1.0 + rng.uniform(-1e-4, 1e-4, n)
    y2 = rng.uniform(-1, 1, n) * 1e6 * rng.random(n)
    x3 = rng.uniform(0, 4, n)
    y3 = rng.uniform(-10, 10, n)
    x = np.concatenate([x1, x2, x3]).astype(np.float32)
    y = np.concatenate([y1, y2, y3]).astype(np.float32)
    got = orc.math2_32("pow", x, y).astype(np.float64)
    with np.errstate(all="ignore"):
        ref = np.power(x.astype(np.float64), y.astype(np.float64))
    fin = np.abs(ref) <= np.finfo(np.float32).max
    assert np.all(np.isinf(got[~fin]))
    sp = np.abs(np.spacing(ref[fin].astype(np.float32))).astype(np.float64)
    err = np.abs(got[fin] - ref[fin]) / np.maximum(sp, 2.0 ** -149)
    assert np.max(err) <= 0.8, (np.max(err), x[fin][np.argmax(err)], y[fin][np.argmax(err)])
    idx = np.flatnonzero(fin)[:: int(fin.sum()) // 2000][:2000]
    with mp.workdps(40):
        want = np.array([float(mp.power(mp.mpf(float(x[i])), mp.mpf(float(y[i])))) for i in idx])
    spm = np.abs(np.spacing(want.astype(np.float32))).astype(np.float64)
    assert np.max(np.abs(got[idx] - want) / np.maximum(spm, 2.0 ** -149)) <= 0.8


def test_float_pow_specials_match_the_double_routine(orc):
    """C99 special cases (signed zeros, infinities, NaN, odd/even integer exponents, negative bases) and
    exact powers: the float form equals the double routine rounded to float, bit for bit."""
    v = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0, 2.5, -2.5, 1e30, -1e30,
         1e-30, 16777216.0, -16777217.0, 8.0, 0.25]
    x, y = np.meshgrid(np.array(v, np.float32), np.array(v, np.float32))
    x, y = x.ravel(), y.ravel()
    got = orc.math2_32("pow", x, y)
    with np.errstate(over="ignore"):
        want = orc.math2_64("pow", x.astype(np.float64), y.astype(np.float64)).astype(np.float32)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert np.all(same), list(zip(x[~same], y[~same], got[~same], want[~same]))[:6]
