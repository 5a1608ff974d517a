"""GPU: seeded random expression trees through the tape (runtime-compiled kernels) against a numpy
evaluation of the same tree in the same order, with the oracle's math functions: bit for bit (NaN payloads
aside), in double and in float.

The trees mix the arithmetic operators (a scalar divisor is a product by its reciprocal, a negation a product
by -1: expr/Operations.h:25-71), the elementwise functions of expr/nodes/Elementwise.h and the binary
min/max/pow forms, over three leaves and a few constants; trees that exceed one program's limits are drawn
again."""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu
CONSTS = [0.5, 2.0, -1.25, 3.0, 0.1]


def same_bits(got, want):
    u = np.uint64 if got.dtype == np.float64 else np.uint32
    eq = (got.view(u) == want.view(u)) | (np.isnan(got) & np.isnan(want))
    return bool(np.all(eq)), np.flatnonzero(~eq)[:4]


class Gen:
    """Random (expression, numpy evaluator) pairs."""

    def __init__(self, orc, leaves, host, dtype, seed):
        self.orc, self.leaves, self.host, self.dt = orc, leaves, host, dtype
        self.r = np.random.default_rng(seed)
        f32 = dtype == np.float32
        self.m1 = orc.math32 if f32 else orc.math64
        self.m2 = orc.math2_32 if f32 else orc.math2_64

    def c(self, v):
        return self.dt.type(v)

    def unary(self, name, x):
        return self.m1(name, np.ascontiguousarray(x))

    def tree(self, depth):
        r = self.r
        if depth == 0 or r.random() < 0.25:
            if r.random() < 0.8:
                k = int(r.integers(0, len(self.leaves)))
                return E.Leaf(self.leaves[k]), (lambda k=k: self.host[k])
            v = float(r.choice(CONSTS))
            return E.Const(v), (lambda v=v: np.full_like(self.host[0], self.c(v)))
        kind = r.random()
        if kind < 0.4:
            ex, f = self.tree(depth - 1)
            op = str(r.choice(["sin", "cos", "exp", "tanh", "atan", "tan", "sign", "abs", "neg", "sqrtabs", "logabs",
                               "recip"]))
            if op == "neg":
                return -ex, (lambda: f() * self.c(-1.0))
            if op == "abs":
                return E.abs_(ex), (lambda: np.abs(f()))
            if op == "sqrtabs":
                return E.sqrt(E.abs_(ex)), (lambda: np.sqrt(np.abs(f())))
            if op == "logabs":
                return E.log(E.abs_(ex) + 0.5), (lambda: self.unary("log", np.abs(f()) + self.c(0.5)))
            if op == "recip":
                s = float(r.choice([2.0, 3.0, -0.5]))
                rc = (np.float32(1) / np.float32(s)) if self.dt == np.float32 else 1.0 / s
                return ex / s, (lambda: f() * self.c(rc))
            fn = {"sin": E.sin, "cos": E.cos, "exp": E.exp, "tanh": E.tanh, "atan": E.atan, "tan": E.tan,
                  "sign": E.sign}[op]
            return fn(ex), (lambda: self.unary(op, f()))
        ex1, f1 = self.tree(depth - 1)
        ex2, f2 = self.tree(depth - 1)
        op = str(r.choice(["+", "-", "*", "/", "min", "max", "powabs"]))
        if op == "+":
            return ex1 + ex2, (lambda: f1() + f2())
        if op == "-":
            return ex1 - ex2, (lambda: f1() - f2())
        if op == "*":
            return ex1 * ex2, (lambda: f1() * f2())
        if op == "/":
            return E.div(ex1, ex2), (lambda: f1() / f2())
        if op == "min":
            return E.minimum(ex1, ex2), (lambda: self.m2("fmin", f1(), f2()))
        if op == "max":
            return E.maximum(ex1, ex2), (lambda: self.m2("fmax", f1(), f2()))
        e = float(r.choice([0.5, 1.5, 2.0, 3.0]))
        return E.pow_(E.abs_(ex1), e), (lambda: self.m2("pow", np.abs(f1()), np.full_like(self.host[0], self.c(e))))


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_random_trees_equal_numpy_and_oracle(dev, orc, dt):
    npdt = np.float64 if dt == torch.float64 else np.float32
    n = 20_001
    r = np.random.default_rng(21)
    host = [r.uniform(-2, 2, n).astype(npdt) for _ in range(3)]
    leaves = []
    for h in host:
        a = ae.Array(n, 1, dt)
        a.data[0, :n].copy_(torch.from_numpy(h).to("cuda"))
        leaves.append(a)
    gen = Gen(orc, leaves, host, np.dtype(npdt), 7 if dt == torch.float64 else 8)
    d = ae.Array(n, 1, dt)
    checked = 0
    while checked < 60:
        ex, f = gen.tree(4)
        try:
            E.lower(d, ex)
        except ValueError:
            continue  # beyond one program's limits: draw again
        E.assign(d, ex)
        torch.cuda.synchronize()
        with np.errstate(all="ignore"):
            want = np.ascontiguousarray(f(), dtype=npdt)
        got = d.data[0, :n].cpu().numpy()
        ok, bad = same_bits(got, want)
        assert ok, (checked, bad, got[bad], want[bad])
        checked += 1


BIG_CONSTS = [0.5, 2.0, -1.25, 3.0, 0.1, 0.75, -0.3, 1.5, 2.5, -2.0, 0.125, 4.0, -0.7, 0.9, 1.1, -1.5, 0.05, 0.6,
              -0.45, 1.75]


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_trees_beyond_one_launch_equal_numpy_and_oracle(dev, orc, dt):
    """Trees of 200+ nodes over 24 arrays and 20 distinct constants -- beyond one launch's 16 loads, 64
    instructions, stack of 8, 16 leaves and 16 constants -- are cut into pieces through temporary planes
    (aeth_eval_tape_ex) and still give numpy's bits with the oracle's math; their fused sums equal the
    standalone reduction of the output."""
    global CONSTS
    npdt = np.float64 if dt == torch.float64 else np.float32
    n = 20_001
    r = np.random.default_rng(31)
    host = [r.uniform(-2, 2, n).astype(npdt) for _ in range(24)]
    leaves = []
    for h in host:
        a = ae.Array(n, 1, dt)
        a.data[0, :n].copy_(torch.from_numpy(h).to("cuda"))
        leaves.append(a)
    saved, CONSTS = CONSTS, BIG_CONSTS
    try:
        gen = Gen(orc, leaves, host, np.dtype(npdt), 41 if dt == torch.float64 else 42)
        d = ae.Array(n, 1, dt)
        checked = 0
        while checked < 6:
            ex, f = gen.tree(9)
            progs, _keep, lv, cs, _ = E.lower_ex(d, ex)
            if progs[0].n_insns < 200:
                continue
            with pytest.raises(ValueError):
                E.lower(d, ex)  # beyond one launch
            sums = E.assign(d, ex, ae.SUM_NEUMAIER if dt == torch.float64 else ae.SUM_FLOATFLOAT)
            torch.cuda.synchronize()
            with np.errstate(all="ignore"):
                want = np.ascontiguousarray(f(), dtype=npdt)
            got = d.data[0, :n].cpu().numpy()
            ok, bad = same_bits(got, want)
            assert ok, (checked, progs[0].n_insns, bad, got[bad], want[bad])
            again = ae.reduce(d, ae.SUM_NEUMAIER if dt == torch.float64 else ae.SUM_FLOATFLOAT)
            assert torch.equal(sums, again) or (torch.isnan(sums).any() and torch.isnan(again).any())
            checked += 1
    finally:
        CONSTS = saved


def test_many_components_and_arrays(dev):
    """A 12-component destination (beyond one launch's 8 programs) of per-component sums over 20 arrays,
    each component reading others' planes: the same as 12 separate assignments."""
    n = 5003
    srcs = [ae.Array(n, 12, torch.float64) for _ in range(20)]
    for k, s in enumerate(srcs):
        ae.array_fill_uniform(s, 100 + k, -1.0, 1.0)
    expr = E.Leaf(srcs[0])
    for k, s in enumerate(srcs[1:]):
        expr = expr + E.Leaf(s) * (0.5 + 0.01 * k)
    d = ae.Array(n, 12, torch.float64)
    E.assign(d, expr)
    torch.cuda.synchronize()
    want = srcs[0].data[:, :n].clone()
    for k, s in enumerate(srcs[1:]):
        c = torch.tensor(0.5 + 0.01 * k, dtype=torch.float64, device="cuda")
        want = want + s.data[:, :n] * c
    assert torch.equal(d.data[:, :n], want)
