"""GPU: the rest of the elementwise set (expr/Elementwise.h, expr/Operations.h) and the per-sample members
(expr/Expression.h, expr/Reduce.h, expr/nodes/Quaternion.h, expr/nodes/Structural.h) through the tape.

Math functions are compared with the oracle bit for bit (NaN payloads aside), in double and in float, on
inputs that cover each function's range and its special values; the same holds on the interpreter path.
The quaternion, fold, norm and slice members are compared bit for bit with numpy restatements that follow
the reference's formulas term by term, and the C++ front end's device results with its own per-sample host
spelling.
"""
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

import aether_amd as ae
import members_ref as M
from aether_amd import build as B
from aether_amd import expr as E

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPECIAL = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 5e-324, -5e-324, 2.2250738585072014e-308]


def load(x, dt):
    """Array of n samples whose planes are the rows of x (ncomp x n)."""
    x = np.atleast_2d(x)
    a = ae.Array(x.shape[1], x.shape[0], dt)
    a.data[:, : x.shape[1]].copy_(torch.from_numpy(np.ascontiguousarray(x)).to(a.data.device))
    return a


def host(a):
    torch.cuda.synchronize()
    return a.data[:, : a.n].cpu().numpy()


def same_bits(got, want):
    got, want = np.asarray(got), np.asarray(want)
    u = np.uint64 if got.dtype == np.float64 else np.uint32
    eq = (got.view(u) == want.view(u)) | (np.isnan(got) & np.isnan(want))
    return bool(np.all(eq)), np.flatnonzero(~eq.reshape(-1))[:5]


def unary_inputs(fn, f32, seed=1):
    r = np.random.default_rng(seed)
    m = 20_000
    top = 30 if f32 else 300
    if fn in ("asin", "acos"):
        x = np.concatenate([r.uniform(-1, 1, m), 1 - 10 ** r.uniform(-16, -1, m // 4),
                            -1 + 10 ** r.uniform(-16, -1, m // 4), [1.5, -2.0, 1.0000000000000002]])
    elif fn == "tan":
        x = np.concatenate([r.uniform(-2, 2, m), 10 ** r.uniform(-9, 7 if not f32 else 5, m) * r.choice([-1, 1], m)])
    elif fn == "tanh":
        x = np.concatenate([r.uniform(-3, 3, m), 10 ** r.uniform(-12, 1.5, m) * r.choice([-1, 1], m)])
    elif fn == "rsqrt":
        x = np.concatenate([10 ** r.uniform(-top, top, m), [-1.0, -0.5]])
    else:  # atan, sign
        x = np.concatenate([r.uniform(-2, 2, m), 10 ** r.uniform(-top + 10, top - 10, m) * r.choice([-1, 1], m)])
    x = np.concatenate([x, SPECIAL])
    return x.astype(np.float32) if f32 else x


UNARY = {"tan": E.tan, "atan": E.atan, "asin": E.asin, "acos": E.acos, "tanh": E.tanh, "rsqrt": E.rsqrt,
         "sign": E.sign}


@pytest.mark.parametrize("fn", list(UNARY))
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_unary_equals_oracle(dev, orc, fn, dt):
    f32 = dt == torch.float32
    x = unary_inputs(fn, f32)
    a, d = load(x, dt), ae.Array(x.size, 1, dt)
    E.assign(d, UNARY[fn](a))
    want = orc.math32(fn, x) if f32 else orc.math64(fn, x)
    ok, bad = same_bits(host(d)[0], want)
    assert ok, (fn, x[bad], host(d)[0][bad], want[bad])


def binary_inputs(f32, seed=2):
    r = np.random.default_rng(seed)
    m = 15_000
    top = 20 if f32 else 300
    ybig = 1e4 if f32 else 1e8
    x = np.concatenate([10 ** r.uniform(-top, top, m), r.uniform(0, 4, m), 1 + r.uniform(-1e-6, 1e-6, m),
                        -r.integers(1, 20, m).astype(float) * r.uniform(0.5, 2, m)])
    y = np.concatenate([r.uniform(-1, 1, m), r.uniform(-60, 60, m), r.uniform(-ybig, ybig, m),
                        r.integers(-30, 30, m).astype(float)])
    s = np.array(SPECIAL + [0.5, -0.5, 2.0, -3.0, 2.0 ** 53])
    sx, sy = np.meshgrid(s, s)
    x, y = np.concatenate([x, sx.ravel()]), np.concatenate([y, sy.ravel()])
    if f32:
        with np.errstate(over="ignore"):
            return x.astype(np.float32), y.astype(np.float32)
    return x, y


BINARY = {"pow": E.pow_, "fmin": E.minimum, "fmax": E.maximum}


@pytest.mark.parametrize("fn", list(BINARY) + ["selgt"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_binary_equals_oracle(dev, orc, fn, dt):
    f32 = dt == torch.float32
    x, y = binary_inputs(f32)
    a, b, d = load(x, dt), load(y, dt), ae.Array(x.size, 1, dt)
    if fn == "selgt":
        # the ordered max has no free function of its own: it is maxNorm's combining step; a one-component
        # maxNorm is T(x_0) = selgt(-x_0, x_0) and a two-component one selgt(T(x_1), T(x_0))
        E.assign(d, E.Leaf(a).maxNorm())
        want = orc.math2_32(fn, -x, x) if f32 else orc.math2_64(fn, -x, x)
    else:
        E.assign(d, BINARY[fn](a, b))
        want = orc.math2_32(fn, x, y) if f32 else orc.math2_64(fn, x, y)
    ok, bad = same_bits(host(d)[0], want)
    assert ok, (fn, x[bad], y[bad], host(d)[0][bad], want[bad])


def test_scalar_forms_and_clamp(dev, orc):
    r = np.random.default_rng(3)
    x = np.concatenate([r.uniform(-2, 2, 30_000), SPECIAL])
    a, d = load(x, torch.float64), ae.Array(x.size, 1, torch.float64)
    E.assign(d, E.clamp(a, -0.5, 0.75) + E.pow_(E.abs_(a), 1.5) - E.div(a, E.Leaf(a) + 3.0))
    with np.errstate(all="ignore"):
        want = (np.fmin(np.fmax(x, -0.5), 0.75) + orc.math2_64("pow", np.abs(x), np.full_like(x, 1.5))) - x / (x + 3.0)
    ok, bad = same_bits(host(d)[0], want)
    assert ok, (x[bad], host(d)[0][bad], want[bad])


from members_ref import dot, fold_max, fold_sum, hamilton, rotate  # noqa: E402,F401


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_members_equal_numpy_restatement(dev, dt):
    npdt = np.float64 if dt == torch.float64 else np.float32
    r = np.random.default_rng(4)
    n = 70_001
    qh, rh, vh = (r.uniform(-1, 1, (k, n)).astype(npdt) for k in (4, 4, 3))
    q, rr, v = load(qh, dt), load(rh, dt), load(vh, dt)
    one = npdt(1)
    L = E.Leaf
    cases = [
        (4, L(q).quatMul(rr), hamilton(qh, rh)),
        (4, L(q).quatConj(), np.stack([qh[0], -qh[1], -qh[2], -qh[3]])),
        (4, L(q).quatReciprocal(), np.stack([qh[0], -qh[1], -qh[2], -qh[3]]) * (one / dot(qh, qh))),
        (3, L(q).quatRotate(v), rotate(qh, vh)),
        (4, L(v).asPureQuaternion(), np.stack([np.zeros(n, npdt), vh[0], vh[1], vh[2]])),
        (3, L(q).asBack3DVector() + L(q).head(3), qh[1:] + qh[:3]),
        (2, L(q).segment(1, 2) - L(v).tail(2), qh[1:3] - vh[1:]),
        (1, L(q).maxNorm(), fold_max(qh)[None]),
        (1, L(v).sum(), fold_sum(vh)[None]),
        (1, L(v).rNorm() + L(v).rSquaredNorm(), (one / np.sqrt(dot(vh, vh)) + one / dot(vh, vh))[None]),
        (1, L(v).cubedNorm() - L(v).rCubedNorm(),
         (dot(vh, vh) * np.sqrt(dot(vh, vh)) - one / (dot(vh, vh) * np.sqrt(dot(vh, vh))))[None]),
        (3, L(v).unitVector(), vh * (one / np.sqrt(dot(vh, vh)))),
        (3, L(v).cwiseMul(L(q).head(3)), vh * qh[:3]),
    ]
    for k, (nc, ex, want) in enumerate(cases):
        d = ae.Array(n, nc, dt)
        E.assign(d, ex)
        ok, bad = same_bits(host(d), want.astype(npdt))
        assert ok, (k, bad)


def test_rotation_by_a_unit_quaternion_is_a_rotation(dev):
    """Size-independent check at 2^22 samples: rotating by q then by q's conjugate gives the vector back
    to a few ulps, and the length is kept."""
    n = 1 << 22
    g = ae.Generator.from_seed(9)
    raw, q, v = ae.Array(n, 4, torch.float64), ae.Array(n, 4, torch.float64), ae.Array(n, 3, torch.float64)
    E.assign(raw, E.uniform(g, 4, -1.0, 1.0))
    E.assign(q, E.Leaf(raw).unitVector())
    E.assign(v, E.uniform(g.substream(1), 3, -1.0, 1.0))
    w, back = ae.Array(n, 3, torch.float64), ae.Array(n, 3, torch.float64)
    E.assign(w, E.Leaf(q).quatRotate(v))
    E.assign(back, E.Leaf(q).quatConj().quatRotate(w))
    vv, ww, bb = host(v), host(w), host(back)
    assert np.max(np.abs(bb - vv)) < 64 * np.finfo(np.float64).eps
    assert np.max(np.abs(np.linalg.norm(ww, axis=0) - np.linalg.norm(vv, axis=0))) < 64 * np.finfo(np.float64).eps


INTERP_SCRIPT = r"""
import numpy as np, torch
import aether_amd as ae
from aether_amd import expr as E
from oracle import pyoracle as orc
r = np.random.default_rng(5)
x = np.concatenate([r.uniform(-1, 1, 5000), 10 ** r.uniform(-200, 200, 5000), [0.0, -0.0, np.inf, -np.inf, np.nan]])
y = np.concatenate([r.uniform(-40, 40, 10000), [2.0, -1.0, 0.5, np.inf, np.nan]])
def load(v):
    a = ae.Array(v.size, 1, torch.float64); a.data[0, :v.size].copy_(torch.from_numpy(v).cuda()); return a
a, b = load(x), load(y)
for name, ex, want in [("tan", E.tan(a), orc.math64("tan", x)), ("atan", E.atan(a), orc.math64("atan", x)),
                       ("asin", E.asin(a), orc.math64("asin", x)), ("acos", E.acos(a), orc.math64("acos", x)),
                       ("tanh", E.tanh(a), orc.math64("tanh", x)), ("sign", E.sign(a), orc.math64("sign", x)),
                       ("pow", E.pow_(E.abs_(a), b), orc.math2_64("pow", np.abs(x), y)),
                       ("fmin", E.minimum(a, b), orc.math2_64("fmin", x, y)),
                       ("fmax", E.maximum(a, b), orc.math2_64("fmax", x, y))]:
    d = ae.Array(x.size, 1, torch.float64)
    E.assign(d, ex)
    torch.cuda.synchronize()
    got = d.data[0, :x.size].cpu().numpy()
    assert np.all((got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))), name
print("interpreter ok")
"""


def test_interpreter_path_same_bits(dev):
    env = dict(os.environ, AETHER_TAPE_INTERPRET="1")
    r = subprocess.run([sys.executable, "-c", INTERP_SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and "interpreter ok" in r.stdout, r.stderr[-2000:]


def test_cpp_members_device_equal_host_and_restatement(dev, orc):
    n = 3000
    assert os.path.exists(B.API_CHECK), "api_check not built (run the package build)"
    out = subprocess.run([B.API_CHECK, "ext", str(n), "17"], capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    flat = np.frombuffer(out.stdout, dtype=np.float64)
    sizes = [4, 4, 3, 4, 3, 1, 3, 4, 3, 1, 3, 3]
    assert flat.size == sum(sizes) * n
    parts, off = [], 0
    for s in sizes:
        parts.append(flat[off: off + s * n].reshape(s, n))
        off += s * n
    q, r, a, o1, o2, o3, o4, h1, h2, h3, h4, o5 = parts
    # the device tape and the host spelling agree bit for bit
    for dv, hv in ((o1, h1), (o2, h2), (o3, h3), (o4, h4)):
        ok, bad = same_bits(dv, hv)
        assert ok, bad
    # and both follow the reference's formulas
    conj = np.stack([q[0], -q[1], -q[2], -q[3]])
    recip = conj * (1.0 / dot(q, q))
    assert same_bits(o1, (hamilton(q, r) + conj) - np.fmax(recip, -2.0))[0]
    assert same_bits(o2, (rotate(q, a) + q[:3] * fold_sum(a)) - a)[0]
    nn = dot(a, a)
    cub = nn * np.sqrt(nn)
    assert same_bits(o3[0], (fold_max(q) + fold_sum(a)) - 1.0 / cub)[0]
    want4 = (((a * (1.0 / np.sqrt(nn)) + np.fmin(a, 0.25)) - np.fmin(np.fmax(a, -0.5), 0.5)) + a / q[1:]) + np.sign(a)
    assert same_bits(o4, want4)[0]
    assert same_bits(o5, orc.math64("tan", a.ravel()).reshape(a.shape) +
                     orc.math2_64("pow", np.abs(a).ravel(), np.full(a.size, 1.5)).reshape(a.shape))[0]


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_members_equal_reference_bits(dev, orc, dt):
    """The reference's own outputs (ref_members_f64/f32.npz, minted from its CPU path): every arithmetic member
    bit for bit; the transcendental ones equal to the oracle bit for bit and within the reference's own
    accuracy of its values."""
    z = M.load_members(np.float64 if dt == torch.float64 else np.float32)
    n = int(z["n"])
    sq, sr, sv = (int(x) for x in z["seeds"])
    q, r, v = ae.Array(n, 4, dt), ae.Array(n, 4, dt), ae.Array(n, 3, dt)
    for arr, seed in ((q, sq), (r, sr), (v, sv)):
        ae.array_fill_uniform(arr, seed, -1.0, 1.0)
    L = E.Leaf
    b3 = L(q).asBack3DVector()
    cases = {
        "quat_mul": L(q).quatMul(r), "quat_conj": L(q).quatConj(), "quat_reciprocal": L(q).quatReciprocal(),
        "as_pure_quaternion": L(v).asPureQuaternion(), "quat_rotate": L(q).quatRotate(v),
        "unit_vector": L(v).unitVector(), "as_back3": b3, "head3": L(q).head(3), "min_025": E.minimum(v, 0.25),
        "max_m025": E.maximum(v, -0.25), "clamp_05": E.clamp(v, -0.5, 0.5), "sign": E.sign(v),
        "div_back3": E.div(v, b3), "mul_head3": L(v).cwiseMul(L(q).head(3)), "min_back3": E.minimum(v, b3),
        "max_back3": E.maximum(v, b3), "tan": E.tan(v), "atan": E.atan(v), "asin": E.asin(v), "acos": E.acos(v),
        "tanh": E.tanh(v), "rsqrt_abs": E.rsqrt(E.abs_(v)), "pow_abs_15": E.pow_(E.abs_(v), 1.5),
        "pow_abs_back3": E.pow_(E.abs_(v), b3), "max_norm": L(q).maxNorm(), "sum": L(v).sum(),
        "norm": L(v).norm(), "squared_norm": L(v).squaredNorm(), "r_norm": L(v).rNorm(),
        "r_squared_norm": L(v).rSquaredNorm(), "cubed_norm": L(v).cubedNorm(), "r_cubed_norm": L(v).rCubedNorm(),
    }
    hq, hv = host(q), host(v)
    trans = M.transcendental(orc, hq, hv)
    for name, ex in cases.items():
        want = z[name]
        d = ae.Array(n, want.shape[0], dt)
        E.assign(d, ex)
        got = host(d)
        if name in trans:
            assert same_bits(got, trans[name])[0], name
            assert np.max(M.ulps(got, want)) <= (2.0 if name == "tanh" else 1.0), name
        else:
            ok, bad = same_bits(got, want)
            assert ok, (name, bad)
