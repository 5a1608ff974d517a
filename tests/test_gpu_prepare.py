"""GPU: aeth_tape_prepare compiles and loads every kernel an assignment would launch, and runs nothing.

After prepare the assignment compiles nothing more (the runtime compiler's counters), prepare leaves the
destination untouched, and the assignment's values and sums are the ones it gives unprepared. Covers a tree
that fits one launch, with and without sums, and a tree of 200+ nodes that is cut into pieces."""
import ctypes as C

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def jit_counts():
    c, h, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
    ae._lib.call("aeth_tape_jit_stats", C.byref(c), C.byref(h), C.byref(f))
    return c.value, h.value


def inputs(n, dt, seeds=(71, 72, 73)):
    out = []
    for s in seeds:
        a = ae.Array(n, 3, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        out.append(a)
    return out


def big_tree(leaves, depth, rng):
    """A random tree of sums, products and sines over single-component leaves (several hundred nodes)."""
    if depth == 0:
        return leaves[int(rng.integers(len(leaves)))] * float(rng.uniform(0.5, 1.5))
    k = int(rng.integers(3))
    a, b = big_tree(leaves, depth - 1, rng), big_tree(leaves, depth - 1, rng)
    if k == 0:
        return a + b
    if k == 1:
        return a * E.sin(b)
    return a - b * 0.75


@pytest.mark.parametrize("dt,mode", [(torch.float64, ae.SUM_NEUMAIER), (torch.float64, ae.SUM_NONE),
                                     (torch.float32, ae.SUM_FLOATFLOAT)])
def test_prepare_then_assign(dev, dt, mode):
    n = 100_003
    a, b, c = inputs(n, dt)
    # a structure no other test compiles: its own constants make it new to this process's cache
    tree = a * 1.25 + b * c - E.tanh(a * 0.5) * E.cos(c) + 0.125
    ref = ae.Array(n, 3, dt)
    want_sums = E.assign(ref, tree, mode)
    torch.cuda.synchronize()
    d = ae.Array(n, 3, dt)
    d.data.fill_(7.0)
    torch.cuda.synchronize()
    E.prepare(d, tree * 1.0, mode)  # one more node: a structure the reference assignment did not compile
    torch.cuda.synchronize()
    assert torch.all(d.data == 7.0), "prepare wrote the destination"
    compiled, _ = jit_counts()
    sums = E.assign(d, tree * 1.0, mode)
    torch.cuda.synchronize()
    assert jit_counts()[0] == compiled, "the assignment compiled after prepare"
    ref2 = ae.Array(n, 3, dt)
    want2 = E.assign(ref2, tree * 1.0, mode)
    torch.cuda.synchronize()
    assert torch.equal(d.data[:, :n], ref2.data[:, :n])
    if mode != ae.SUM_NONE:
        assert torch.equal(sums, want2)
        assert want_sums is not None


def test_prepare_tree_cut_into_pieces(dev):
    n = 20_001
    rng = np.random.default_rng(5)
    leaves = []
    for s in range(12):
        x = ae.Array(n, 1, torch.float64)
        ae.array_fill_uniform(x, 900 + s, -1.0, 1.0)
        leaves.append(x)
    tree = big_tree(leaves, 7, rng)
    d = ae.Array(n, 1, torch.float64)
    progs, _keep, _lv, _cs, _ = E.lower_ex(d, tree)
    assert progs[0].n_insns > 64, "the tree must be beyond one launch"
    d.data.fill_(-3.0)
    E.prepare(d, tree, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert torch.all(d.data == -3.0)
    compiled, _ = jit_counts()
    sums = E.assign(d, tree, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert jit_counts()[0] == compiled
    again = ae.reduce(d, ae.SUM_NEUMAIER)
    assert torch.equal(sums, again)
