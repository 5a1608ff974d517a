"""Expression trees over device Arrays, lowered to one HIP kernel per assignment.

The Python spelling of aether's lazy assignment (expr/Operations.h:25-71,
expr/nodes/Arithmetic.h:33-143, expr/nodes/Elementwise.h:245-533,
expr/nodes/Geometric.h:32-73, expr/Reduce.h:150-190):

    d = a + b * c + sin(a) * exp(-1.0 * (b * b))
    assign(D, d)                       # one kernel, optional fused sums

Each output component becomes one program (aeth_eval_tape): the planes it
reads and a postfix tape in aether's evaluation order. `-x` is `x * -1`
(CWiseScale), `x / s` for a scalar s is `x * (1/s)` with the reciprocal
rounded in the element type (expr/Operations.h:62-71), folds are
right-nested `e[N-1] + (... + e[0])` (docs/content/numerics_contract.rst:12-24).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import AETH_RNG_DIST_NORMAL, AETH_RNG_DIST_UNIFORM, Array, _dtype_code, stream_handle
from ._lib import call

OP = {"push": 0, "const": 1, "rng": 2, "+": 3, "-": 4, "*": 5, "/": 6, "neg": 7, "sin": 8, "cos": 9, "exp": 10,
      "log": 11, "sqrt": 12, "abs": 13, "sign": 14, "tan": 15, "atan": 16, "asin": 17, "acos": 18, "tanh": 19,
      "fmin": 20, "fmax": 21, "selgt": 22, "pow": 23}
MAX_LOADS, MAX_INSNS, MAX_PROGRAMS, MAX_LEAVES, MAX_CONSTS, MAX_RNGS = 16, 64, 8, 16, 16, 8


class aeth_insn(C.Structure):
    _fields_ = [("op", C.c_uint8), ("arg", C.c_uint8)]


class aeth_program(C.Structure):
    _fields_ = [("n_loads", C.c_uint8), ("n_insns", C.c_uint8), ("out_leaf", C.c_uint8), ("out_comp", C.c_uint8),
                ("load_leaf", C.c_uint8 * MAX_LOADS), ("load_comp", C.c_uint8 * MAX_LOADS),
                ("insns", aeth_insn * MAX_INSNS)]


class aeth_insn_ex(C.Structure):
    _fields_ = [("op", C.c_uint8), ("pad", C.c_uint8 * 3), ("arg", C.c_uint32)]


class aeth_program_ex(C.Structure):
    _fields_ = [("n_loads", C.c_uint32), ("n_insns", C.c_uint32), ("out_leaf", C.c_uint32), ("out_comp", C.c_uint32),
                ("load_leaf", C.POINTER(C.c_uint32)), ("load_comp", C.POINTER(C.c_uint32)),
                ("insns", C.POINTER(aeth_insn_ex))]


class aeth_rng_leaf(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("p0", C.c_double), ("p1", C.c_double), ("comp", C.c_uint32),
                ("dist", C.c_uint32)]


def _wrap(x) -> "Expr":
    if isinstance(x, Expr):
        return x
    if isinstance(x, Array):
        return Leaf(x)
    if isinstance(x, (int, float)):
        return Const(float(x))
    raise TypeError(f"cannot use {type(x)} in an expression")


class Expr:
    ncomp = 1

    def __add__(self, o): return Bin("+", self, _wrap(o))
    def __radd__(self, o): return Bin("+", _wrap(o), self)
    def __sub__(self, o): return Bin("-", self, _wrap(o))
    def __rsub__(self, o): return Bin("-", _wrap(o), self)
    def __mul__(self, o): return Bin("*", self, _wrap(o))
    def __rmul__(self, o): return Bin("*", _wrap(o), self)

    def __truediv__(self, o):
        if isinstance(o, (int, float)):
            return Recip(self, float(o))  # x * (1/s): expr/Operations.h:62-71
        return Bin("/", self, _wrap(o))

    def __neg__(self): return Bin("*", self, Const(-1.0))  # numerics_contract.rst:66

    def dot(self, o): return Dot(self, _wrap(o))
    def cross(self, o): return Cross(self, _wrap(o))
    def squared_norm(self): return Dot(self, self)
    def norm(self): return Un("sqrt", Dot(self, self))
    # expr/Expression.h:196-242: the reciprocal / cubed norms, the ordered max-abs and the sum
    def r_norm(self): return Bin("/", Const(1.0), self.norm())
    def r_squared_norm(self): return Bin("/", Const(1.0), Dot(self, self))
    def cubed_norm(self): return Bin("*", Dot(self, self), self.norm())
    def r_cubed_norm(self): return Bin("/", Const(1.0), self.cubed_norm())
    def max_norm(self): return Fold(self, True)
    def sum(self): return Fold(self, False)
    def unit_vector(self): return Bin("*", self, self.r_norm())
    def cwise_mul(self, o): return Bin("*", self, _wrap(o))
    # expr/nodes/Quaternion.h, expr/nodes/Structural.h
    def quat_mul(self, o): return QuatMul(self, _wrap(o))
    def quat_conj(self): return QuatConj(self)
    def quat_reciprocal(self): return Bin("*", QuatConj(self), self.r_squared_norm())
    def quat_rotate(self, v): return QuatRotate(self, _wrap(v))
    def as_pure_quaternion(self): return PureQuat(self)
    def segment(self, off, length): return Segment(self, off, length)
    def head(self, n): return Segment(self, 0, n)
    def tail(self, n): return Segment(self, self.ncomp - n, n)
    def as_back_3d_vector(self): return Segment(self, self.ncomp - 3, 3)
    # the reference's spellings
    rNorm, rSquaredNorm, cubedNorm, rCubedNorm, maxNorm = r_norm, r_squared_norm, cubed_norm, r_cubed_norm, max_norm
    squaredNorm, unitVector, cwiseMul, quatMul, quatConj = squared_norm, unit_vector, cwise_mul, quat_mul, quat_conj
    quatReciprocal, quatRotate, asPureQuaternion, asBack3DVector = (quat_reciprocal, quat_rotate, as_pure_quaternion,
                                                                     as_back_3d_vector)

    def emit(self, c: int, ctx: "_Ctx") -> None:
        raise NotImplementedError


class Leaf(Expr):
    def __init__(self, arr: Array):
        self.arr = arr
        self.ncomp = arr.ncomp

    def emit(self, c, ctx):
        ctx.push(self.arr, c if self.ncomp > 1 else 0)


class Const(Expr):
    def __init__(self, v: float):
        self.v = v

    def emit(self, c, ctx):
        ctx.const(self.v)


class Recip(Expr):
    def __init__(self, e: Expr, s: float):
        self.e, self.s = e, s
        self.ncomp = e.ncomp

    def emit(self, c, ctx):
        self.e.emit(c, ctx)
        r = (np.float32(1) / np.float32(self.s)) if ctx.dtype == torch.float32 else 1.0 / self.s
        ctx.const(float(r))
        ctx.op("*")


class Random(Expr):
    """A distribution leaf (random/detail/DistributionLeaf.h:111-236)."""

    def __init__(self, seed: int, ncomp: int, dist: int, p0: float, p1: float):
        self.seed, self.ncomp, self.dist, self.p0, self.p1 = seed, ncomp, dist, p0, p1

    def emit(self, c, ctx):
        ctx.rng(self.seed, c, self.dist, self.p0, self.p1)


def uniform(generator, ncomp: int, lo: float, hi: float) -> Random:
    """Generator(s).uniform<Real, ncomp>(lo, hi) (random/Generator.h:86-90)."""
    return Random(generator.uniform_seed(), ncomp, AETH_RNG_DIST_UNIFORM, lo, hi)


def normal(generator, ncomp: int, mean: float, sd: float) -> Random:
    """Generator(s).normal<Real, ncomp>(mean, sd) (random/Generator.h:94-104)."""
    return Random(generator.normal_seed(), ncomp, AETH_RNG_DIST_NORMAL, mean, sd)


# the other distributions' stream salts (random/detail/DistributionLeaf.h:249, 317, 378)
LOGNORMAL_SALT, EXPONENTIAL_SALT, BERNOULLI_SALT = 0xA4093822299F31D0, 0x082EFA98EC4E6C89, 0x452821E638D01377


def lognormal(generator, ncomp: int, mu: float, sigma: float) -> "Expr":
    """Generator(s).lognormal<Real, ncomp>(mu, sigma) (random/detail/DistributionLeaf.h:240-300):
    exp(mu + sigma z), z the standard normal of the lognormal stream."""
    return exp(Random(generator.key ^ LOGNORMAL_SALT, ncomp, AETH_RNG_DIST_NORMAL, mu, sigma))


def exponential(generator, ncomp: int, lam: float) -> "Expr":
    """Generator(s).exponential<Real, ncomp>(lambda) (DistributionLeaf.h:305-365): -log(u) / lambda, a
    division (not a product by 1/lambda)."""
    u = Random(generator.key ^ EXPONENTIAL_SALT, ncomp, AETH_RNG_DIST_UNIFORM, 0.0, 1.0)
    return div(-log(u), Const(float(lam)))


def bernoulli(generator, ncomp: int, p: float) -> "Expr":
    """Generator(s).bernoulli<Real, ncomp>(p) (DistributionLeaf.h:370-430): 1 when u < p, else 0; sign(p - u)
    is +1 exactly when u < p (a difference of two floats is 0 only when they are equal)."""
    u = Random(generator.key ^ BERNOULLI_SALT, ncomp, AETH_RNG_DIST_UNIFORM, 0.0, 1.0)
    return maximum(sign(Const(float(p)) - u), 0.0)


class Bin(Expr):
    def __init__(self, op, l: Expr, r: Expr):
        self.op, self.l, self.r = op, l, r
        self.ncomp = max(l.ncomp, r.ncomp)
        if l.ncomp > 1 and r.ncomp > 1 and l.ncomp != r.ncomp:
            raise ValueError("component counts differ")

    def emit(self, c, ctx):
        self.l.emit(c if self.l.ncomp > 1 else 0, ctx)
        self.r.emit(c if self.r.ncomp > 1 else 0, ctx)
        ctx.op(self.op)


class Un(Expr):
    def __init__(self, fn, e: Expr):
        self.fn, self.e = fn, e
        self.ncomp = e.ncomp

    def emit(self, c, ctx):
        self.e.emit(c, ctx)
        ctx.op(self.fn)


class Dot(Expr):
    """Right-nested fold (x_{N-1}*y_{N-1}) + (... + (x_0*y_0)) (expr/Reduce.h:172-188)."""

    def __init__(self, a: Expr, b: Expr):
        if a.ncomp != b.ncomp:
            raise ValueError("dot: component counts differ")
        self.a, self.b = a, b

    def emit(self, c, ctx):
        n = self.a.ncomp
        # postfix of (t_{n-1} + (t_{n-2} + (... + t_0)))
        for k in range(n - 1, -1, -1):
            self.a.emit(k, ctx)
            self.b.emit(k, ctx)
            ctx.op("*")
        for _ in range(n - 1):
            ctx.op("+")


class Cross(Expr):
    """Geometric.h:55-71: (l1 r2 - l2 r1, l2 r0 - l0 r2, l0 r1 - l1 r0)."""

    ncomp = 3

    def __init__(self, l: Expr, r: Expr):
        if l.ncomp != 3 or r.ncomp != 3:
            raise ValueError("cross needs 3-vectors")
        self.l, self.r = l, r

    def emit(self, c, ctx):
        i, j = [(1, 2), (2, 0), (0, 1)][c]
        self.l.emit(i, ctx)
        self.r.emit(j, ctx)
        ctx.op("*")
        self.l.emit(j, ctx)
        self.r.emit(i, ctx)
        ctx.op("*")
        ctx.op("-")


class Fold(Expr):
    """expr/Reduce.h:150-164: acc = T(x_0); acc = op(T(x_{k+1}), acc). sum: T = x, op = +; max-abs:
    T(x) = x < 0 ? -x : x (as selgt(-x, x)), op = the ordered max selgt(a, b) = a > b ? a : b. All
    terms are pushed first (highest component first) so each op's left operand is the new term."""

    ncomp = 1

    def __init__(self, e: Expr, is_max: bool):
        self.e, self.is_max = e, is_max

    def emit(self, c, ctx):
        n = self.e.ncomp
        for k in range(n - 1, -1, -1):
            self.e.emit(k, ctx)
            if self.is_max:
                ctx.op("neg")
                self.e.emit(k, ctx)
                ctx.op("selgt")
        for _ in range(n - 1):
            ctx.op("selgt" if self.is_max else "+")


class QuatMul(Expr):
    """Quaternion.h:35-80, each component's sum left to right as written there."""

    ncomp = 4
    _B = [(0, 1, 2, 3), (1, 0, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]
    _S = [("-", "-", "-"), ("+", "+", "-"), ("-", "+", "+"), ("+", "-", "+")]

    def __init__(self, l: Expr, r: Expr):
        if l.ncomp != 4 or r.ncomp != 4:
            raise ValueError("quat_mul needs quaternions")
        self.l, self.r = l, r

    def emit(self, c, ctx):
        for t in range(4):
            self.l.emit(t, ctx)
            self.r.emit(self._B[c][t], ctx)
            ctx.op("*")
            if t:
                ctx.op(self._S[c][t - 1])


class QuatConj(Expr):
    """Quaternion.h:83-110: (x0, -x1, -x2, -x3), true negation."""

    ncomp = 4

    def __init__(self, e: Expr):
        if e.ncomp != 4:
            raise ValueError("quat_conj needs a quaternion")
        self.e = e

    def emit(self, c, ctx):
        self.e.emit(c, ctx)
        if c:
            ctx.op("neg")


class QuatRotate(Expr):
    """Quaternion.h:117-170: t = 2 (u x v), v' = v + w t + u x t (u the vector part), sums as written."""

    ncomp = 3
    _I, _J = (1, 2, 0), (2, 0, 1)

    def __init__(self, q: Expr, v: Expr):
        if q.ncomp != 4 or v.ncomp != 3:
            raise ValueError("quat_rotate needs a quaternion and a 3-vector")
        self.q, self.v = q, v

    def _t(self, c, ctx):
        i, j = self._I[c], self._J[c]
        ctx.const(2.0)
        self.q.emit(1 + i, ctx)
        self.v.emit(j, ctx)
        ctx.op("*")
        self.q.emit(1 + j, ctx)
        self.v.emit(i, ctx)
        ctx.op("*")
        ctx.op("-")
        ctx.op("*")

    def emit(self, c, ctx):
        i, j = self._I[c], self._J[c]
        self.v.emit(c, ctx)
        self.q.emit(0, ctx)
        self._t(c, ctx)
        ctx.op("*")
        ctx.op("+")
        self.q.emit(1 + i, ctx)
        self._t(j, ctx)
        ctx.op("*")
        self.q.emit(1 + j, ctx)
        self._t(i, ctx)
        ctx.op("*")
        ctx.op("-")
        ctx.op("+")


class PureQuat(Expr):
    """Quaternion.h:173-205: (0, x0, x1, x2)."""

    ncomp = 4

    def __init__(self, e: Expr):
        if e.ncomp != 3:
            raise ValueError("as_pure_quaternion needs a 3-vector")
        self.e = e

    def emit(self, c, ctx):
        if c == 0:
            ctx.const(0.0)
        else:
            self.e.emit(c - 1, ctx)


class Segment(Expr):
    """Structural.h Segment: components off .. off + length - 1."""

    def __init__(self, e: Expr, off: int, length: int):
        if off < 0 or length < 1 or off + length > e.ncomp:
            raise ValueError("segment out of range")
        self.e, self.off, self.ncomp = e, off, length

    def emit(self, c, ctx):
        self.e.emit(self.off + c, ctx)


def sin(e): return Un("sin", _wrap(e))
def cos(e): return Un("cos", _wrap(e))
def exp(e): return Un("exp", _wrap(e))
def log(e): return Un("log", _wrap(e))
def sqrt(e): return Un("sqrt", _wrap(e))
def abs_(e): return Un("abs", _wrap(e))


def sign(e): return Un("sign", _wrap(e))
def tan(e): return Un("tan", _wrap(e))
def atan(e): return Un("atan", _wrap(e))
def asin(e): return Un("asin", _wrap(e))
def acos(e): return Un("acos", _wrap(e))
def tanh(e): return Un("tanh", _wrap(e))
def rsqrt(e): return Bin("/", Const(1.0), Un("sqrt", _wrap(e)))  # math/root.h:74: 1 / sqrt(x)
def minimum(a, b): return Bin("fmin", _wrap(a), _wrap(b))
def maximum(a, b): return Bin("fmax", _wrap(a), _wrap(b))
def clamp(e, lo, hi): return minimum(maximum(e, lo), hi)  # Elementwise.h:219-224
def div(a, b): return Bin("/", _wrap(a), _wrap(b))
def pow_(a, b): return Bin("pow", _wrap(a), _wrap(b))


# aether's names (expr/nodes/Elementwise.h:194-851)
cwiseSin, cwiseCos, cwiseExp, cwiseLog, cwiseSqrt, cwiseAbs = sin, cos, exp, log, sqrt, abs_
cwiseSign, cwiseTan, cwiseAtan, cwiseAsin, cwiseAcos, cwiseTanh, cwiseRsqrt = sign, tan, atan, asin, acos, tanh, rsqrt
cwiseMin, cwiseMax, cwiseDiv, cwisePow = minimum, maximum, div, pow_


class _Ctx:
    def __init__(self, dtype, leaves, consts, rngs):
        self.dtype, self.leaves, self.consts, self.rngs = dtype, leaves, consts, rngs
        self.loads, self.insns = [], []

    def _leaf(self, arr: Array) -> int:
        for k, x in enumerate(self.leaves):
            if x is arr:
                return k
        self.leaves.append(arr)
        return len(self.leaves) - 1

    def push(self, arr, comp):
        key = (self._leaf(arr), comp)
        if key not in self.loads:
            self.loads.append(key)
        self.insns.append((OP["push"], self.loads.index(key)))

    def const(self, v):
        if v not in self.consts:
            self.consts.append(v)
        self.insns.append((OP["const"], self.consts.index(v)))

    def rng(self, seed, comp, dist, p0, p1):
        key = (seed & ((1 << 64) - 1), float(p0), float(p1), comp, dist)
        if key not in self.rngs:
            self.rngs.append(key)
        self.insns.append((OP["rng"], self.rngs.index(key)))

    def op(self, name):
        self.insns.append((OP[name], 0))


def _programs(dest: Array, expr):
    """Per destination component: (loads, insns, out_leaf), with the shared leaf, constant and random-leaf
    tables."""
    expr = _wrap(expr)
    if expr.ncomp not in (1, dest.ncomp):
        raise ValueError("expression and destination component counts differ")
    leaves, consts, rngs, out = [], [], [], []
    for c in range(dest.ncomp):
        ctx = _Ctx(dest.dtype, leaves, consts, rngs)
        expr.emit(c if expr.ncomp > 1 else 0, ctx)
        out.append((ctx.loads, ctx.insns, ctx._leaf(dest)))
    return out, leaves, consts, rngs


def _stack_depth(insns):
    d = peak = 0
    for o, _ in insns:
        if o in (OP["push"], OP["const"], OP["rng"]):
            d += 1
        elif o in _BINARY:
            d -= 1
        peak = max(peak, d)
    return peak


_BINARY = {OP[k] for k in ("+", "-", "*", "/", "fmin", "fmax", "selgt", "pow")}


def lower(dest: Array, expr: Expr):
    """Lower `dest = expr` to the programs of ONE launch (programs, leaves, consts, rngs); ValueError when the
    tree is beyond one launch's limits (assign() still runs it, cut into pieces: lower_ex)."""
    progs_list, leaves, consts, rngs = _programs(dest, expr)
    if (dest.ncomp > MAX_PROGRAMS or len(leaves) > MAX_LEAVES or len(consts) > MAX_CONSTS
            or len(rngs) > MAX_RNGS):
        raise ValueError("too many components, arrays, constants or random leaves for one launch")
    progs = (aeth_program * dest.ncomp)()
    for c, (loads, insns, out_leaf) in enumerate(progs_list):
        if len(loads) > MAX_LOADS or len(insns) > MAX_INSNS or _stack_depth(insns) > 8:
            raise ValueError("expression too large for one program")
        p = progs[c]
        p.n_loads, p.n_insns, p.out_leaf, p.out_comp = len(loads), len(insns), out_leaf, c
        for k, (lf, cp) in enumerate(loads):
            p.load_leaf[k], p.load_comp[k] = lf, cp
        for k, (o, a) in enumerate(insns):
            p.insns[k].op, p.insns[k].arg = o, a
    return progs, leaves, consts, rngs


def lower_ex(dest: Array, expr: Expr):
    """Lower `dest = expr` without size limits (aeth_eval_tape_ex): (programs, keep-alive tables, leaves,
    consts, rngs)."""
    progs_list, leaves, consts, rngs = _programs(dest, expr)
    progs = (aeth_program_ex * dest.ncomp)()
    keep = []
    for c, (loads, insns, out_leaf) in enumerate(progs_list):
        ll = (C.c_uint32 * max(1, len(loads)))(*[lf for lf, _ in loads])
        lc = (C.c_uint32 * max(1, len(loads)))(*[cp for _, cp in loads])
        ins = (aeth_insn_ex * len(insns))()
        for k, (o, a) in enumerate(insns):
            ins[k].op, ins[k].arg = o, a
        keep += [ll, lc, ins]
        p = progs[c]
        p.n_loads, p.n_insns, p.out_leaf, p.out_comp = len(loads), len(insns), out_leaf, c
        p.load_leaf, p.load_comp = C.cast(ll, C.POINTER(C.c_uint32)), C.cast(lc, C.POINTER(C.c_uint32))
        p.insns = C.cast(ins, C.POINTER(aeth_insn_ex))
    return progs, keep, leaves, consts, rngs


class aeth_tape_shape(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("vector_bytes", C.c_uint32), ("waves_per_simd", C.c_uint32), ("ms", C.c_float)]


def _abi_args(dest: Array, expr, unbounded: bool = False):
    if unbounded:
        progs, keep, leaves, consts, rngs = lower_ex(dest, _wrap(expr))
    else:
        progs, leaves, consts, rngs = lower(dest, _wrap(expr))
        keep = None
    for lf in leaves:
        if lf.n != dest.n or lf.dtype != dest.dtype:
            raise ValueError("all arrays of an assignment share sample count and dtype")
    ptrs = (C.c_void_p * len(leaves))(*[lf.ptr() for lf in leaves])
    pitch = (C.c_uint64 * len(leaves))(*[lf.capacity for lf in leaves])
    cst = (C.c_double * max(1, len(consts)))(*consts)
    rg = (aeth_rng_leaf * max(1, len(rngs)))()
    for k, (seed, p0, p1, comp, dist) in enumerate(rngs):
        rg[k].seed, rg[k].p0, rg[k].p1, rg[k].comp, rg[k].dist = seed, p0, p1, comp, dist
    return progs, len(leaves), ptrs, pitch, len(consts), cst, len(rngs), rg, keep


def tune(dest: Array, expr, stream=None, sum_mode: int = -1) -> dict:
    """Autotune the launch shape of `dest = expr` (aeth_tape_tune): every candidate shape (samples per tile,
    bytes per vector access, minimum waves per SIMD) runs on these arrays, the fastest is kept for this
    expression's structure, and `dest` holds the result. With a sum_mode (aeth_tape_tune_sums) only the
    occupancy target varies, the sums' shape stays, and later assignments with that sum mode use the pick.
    Returns {"tile", "vector_bytes", "waves_per_simd", "ms"}."""
    progs, nl, ptrs, pitch, nc, cst, nr, rg, _ = _abi_args(dest, expr)
    out = aeth_tape_shape()
    if sum_mode >= 0:
        sums = torch.empty((dest.ncomp, 2), dtype=dest.dtype, device=dest.data.device)
        call("aeth_tape_tune_sums", _dtype_code(dest.dtype), dest.ncomp, progs, nl, ptrs, pitch, nc, cst, nr, rg,
             dest.n, sum_mode, sums.data_ptr(), stream_handle(stream), C.byref(out))
    else:
        call("aeth_tape_tune", _dtype_code(dest.dtype), dest.ncomp, progs, nl, ptrs, pitch, nc, cst, nr, rg, dest.n,
             stream_handle(stream), C.byref(out))
    return {"tile": out.tile, "vector_bytes": out.vector_bytes, "waves_per_simd": out.waves_per_simd, "ms": out.ms}


def prepare(dest: Array, expr, sum_mode: int = -1, stream=None):
    """Compile and load every kernel `assign(dest, expr, sum_mode)` would launch, without running anything, so
    the first real assignment pays no runtime compile (no counterpart in the reference, whose kernels are
    compiled ahead of time)."""
    progs, nl, ptrs, pitch, nc, cst, nr, rg, keep = _abi_args(dest, expr, unbounded=True)
    call("aeth_tape_prepare", _dtype_code(dest.dtype), dest.ncomp, progs, nl, ptrs, pitch, nc, cst, nr, rg, dest.n,
         sum_mode, stream_handle(stream))
    del keep


def kernel_timing(enable: bool) -> None:
    """Bracket every tape kernel launched from now on (outside stream capture) with two HIP events on its
    stream, or stop doing so; the sums' group tree after each launch is outside the pair."""
    call("aeth_tape_timing", 1 if enable else 0)


def kernel_times() -> list:
    """Milliseconds of each tape kernel launched while timing was on since the last read, in launch order
    (waits for them to finish)."""
    count = C.c_int32()
    call("aeth_tape_timing_read", None, 0, C.byref(count))
    n = count.value
    buf = (C.c_float * max(n, 1))()
    call("aeth_tape_timing_read", buf, n, C.byref(count))
    return [float(buf[k]) for k in range(n)]


def assign(dest: Array, expr, sum_mode: int = -1, sums: torch.Tensor = None, stream=None):
    """`dest = expr` over all samples: ONE kernel when the tree fits one launch, otherwise the tree cut into
    pieces that do (same values); returns (ncomp, 2) sums when sum_mode >= 0."""
    progs, nl, ptrs, pitch, nc, cst, nr, rg, keep = _abi_args(dest, expr, unbounded=True)
    if sum_mode >= 0 and sums is None:
        sums = torch.empty((dest.ncomp, 2), dtype=dest.dtype, device=dest.data.device)
    call("aeth_eval_tape_ex", _dtype_code(dest.dtype), dest.ncomp, progs, nl, ptrs, pitch, nc, cst, nr, rg, dest.n,
         sum_mode, sums.data_ptr() if sums is not None else None, stream_handle(stream))
    del keep
    return sums
