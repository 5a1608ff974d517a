"""Numpy restatements of the reference's per-sample members, term by term in the reference's order, and the
loader of the members fixture minted from the reference's own CPU path (tests/golden/ref_members_f64.npz).
Shared by the oracle tests and the GPU tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hamilton(l, r):
    """expr/nodes/Quaternion.h:66-75, each sum left to right."""
    return np.stack([l[0] * r[0] - l[1] * r[1] - l[2] * r[2] - l[3] * r[3],
                     l[0] * r[1] + l[1] * r[0] + l[2] * r[3] - l[3] * r[2],
                     l[0] * r[2] - l[1] * r[3] + l[2] * r[0] + l[3] * r[1],
                     l[0] * r[3] + l[1] * r[2] - l[2] * r[1] + l[3] * r[0]])


def rotate(q, v):
    """expr/nodes/Quaternion.h:151-166: t = 2 (u x v), v + w t + (u x t)."""
    w, ux, uy, uz = q
    vx, vy, vz = v
    two = q.dtype.type(2)
    tx, ty, tz = two * (uy * vz - uz * vy), two * (uz * vx - ux * vz), two * (ux * vy - uy * vx)
    return np.stack([vx + w * tx + (uy * tz - uz * ty), vy + w * ty + (uz * tx - ux * tz),
                     vz + w * tz + (ux * ty - uy * tx)])


def conj(q):
    """expr/nodes/Quaternion.h:100-107."""
    return np.stack([q[0], -q[1], -q[2], -q[3]])


def dot(x, y):
    """expr/Reduce.h:172-188: x_{N-1} y_{N-1} + (... + x_0 y_0)."""
    acc = x[0] * y[0]
    for k in range(1, len(x)):
        acc = x[k] * y[k] + acc
    return acc


def fold_sum(x):
    """expr/Reduce.h:150-164 with the sum: acc = x_{k+1} + acc."""
    acc = x[0]
    for k in range(1, len(x)):
        acc = x[k] + acc
    return acc


def fold_max(x):
    """expr/Reduce.h:74-130,150-164: T(x) = x < 0 ? -x : x; acc = T(x_{k+1}) > acc ? T(x_{k+1}) : acc."""
    t = lambda v: np.where(v < 0, -v, v)  # noqa: E731
    acc = t(x[0])
    for k in range(1, len(x)):
        acc = np.where(t(x[k]) > acc, t(x[k]), acc)
    return acc


def load_members(dtype=np.float64):
    return np.load(os.path.join(GOLDEN, "ref_members_f64.npz" if dtype == np.float64 else "ref_members_f32.npz"))


def members_inputs(orc, z, dtype=np.float64):
    """q, r (4 planes) and v (3 planes): raw-seed uniforms on (-1, 1) with the fixture's seeds."""
    n = int(z["n"])
    sq, sr, sv = (int(s) for s in z["seeds"])
    f = lambda seed, k: orc.fill_array_uniform(seed, k, n, n, dtype=dtype).reshape(k, n)  # noqa: E731
    return f(sq, 4), f(sr, 4), f(sv, 3)


def restated(q, r, v):
    """Every arithmetic member of the fixture, restated (no transcendental functions)."""
    one = v.dtype.type(1)
    b3 = q[1:]
    nn = dot(v, v)
    with np.errstate(divide="ignore", invalid="ignore"):
        return {
            "quat_mul": hamilton(q, r), "quat_conj": conj(q), "quat_reciprocal": conj(q) * (one / dot(q, q)),
            "as_pure_quaternion": np.stack([np.zeros_like(v[0]), v[0], v[1], v[2]]),
            "quat_rotate": rotate(q, v), "unit_vector": v * (one / np.sqrt(nn)), "as_back3": b3, "head3": q[:3],
            "min_025": np.fmin(v, v.dtype.type(0.25)), "max_m025": np.fmax(v, v.dtype.type(-0.25)),
            "clamp_05": np.fmin(np.fmax(v, v.dtype.type(-0.5)), v.dtype.type(0.5)),
            "sign": np.sign(v), "div_back3": v / b3, "mul_head3": v * q[:3], "min_back3": np.fmin(v, b3),
            "max_back3": np.fmax(v, b3),
            "max_norm": fold_max(q)[None], "sum": fold_sum(v)[None], "norm": np.sqrt(nn)[None],
            "squared_norm": nn[None], "r_norm": (one / np.sqrt(nn))[None], "r_squared_norm": (one / nn)[None],
            "cubed_norm": (nn * np.sqrt(nn))[None], "r_cubed_norm": (one / (nn * np.sqrt(nn)))[None],
        }


# transcendental members: name -> (oracle function, argument builder)
TRANSCENDENTAL = ["tan", "atan", "asin", "acos", "tanh", "rsqrt_abs", "pow_abs_15", "pow_abs_back3"]


def transcendental(orc, q, v):
    """The oracle's values of the transcendental members (double or float, as v is)."""
    av = np.abs(v)
    f32 = v.dtype == np.float32
    m1, m2 = (orc.math32, orc.math2_32) if f32 else (orc.math64, orc.math2_64)
    f = lambda fn, x: m1(fn, x.ravel()).reshape(x.shape)  # noqa: E731
    one = v.dtype.type(1)
    return {"tan": f("tan", v), "atan": f("atan", v), "asin": f("asin", v), "acos": f("acos", v),
            "tanh": f("tanh", v), "rsqrt_abs": one / np.sqrt(av),
            "pow_abs_15": m2("pow", av.ravel(), np.full(av.size, 1.5, v.dtype)).reshape(v.shape),
            "pow_abs_back3": m2("pow", av.ravel(), q[1:].ravel()).reshape(v.shape)}


def ulps(a, b):
    """|a - b| in units in the last place of b, in b's own precision."""
    b = np.asarray(b)
    sp = np.spacing(np.abs(b)).astype(np.float64)
    return np.abs(np.asarray(a, np.float64) - b.astype(np.float64)) / sp
