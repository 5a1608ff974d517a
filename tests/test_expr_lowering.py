"""Host logic: expression trees lower to the tapes aether's evaluation order implies.

CPU only: builds programs for arrays that live on the CPU (no kernel runs).
Checks postfix order, load de-duplication, constant rules (-x is x*-1,
x/s is x*(1/s) rounded in the element type: expr/Operations.h:49-71) and the
right-nested dot fold (expr/Reduce.h:172-188).
"""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

NAMES = {v: k for k, v in E.OP.items()}


def tape(p):
    return [(NAMES[p.insns[k].op], p.insns[k].arg) for k in range(p.n_insns)]


def arr(ncomp=3, dt=torch.float64):
    return ae.Array(10, ncomp, dt, device="cpu")


def test_cfg0_expression_tape():
    a, b, c, d = arr(), arr(), arr(), arr()
    x = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))
    progs, leaves, consts, rngs = E.lower(d, x)
    assert len(progs) == 3 and leaves[:3] == [a, b, c] and leaves[3] is d
    p = progs[1]
    assert p.out_comp == 1 and p.n_loads == 3
    assert [(p.load_leaf[k], p.load_comp[k]) for k in range(3)] == [(0, 1), (1, 1), (2, 1)]
    assert tape(p) == [("push", 0), ("push", 1), ("push", 2), ("*", 0), ("+", 0), ("push", 0), ("sin", 0),
                       ("const", 0), ("push", 1), ("push", 1), ("*", 0), ("*", 0), ("exp", 0), ("*", 0), ("+", 0)]
    assert consts == [-1.0]


def test_negation_and_scalar_division():
    a, d = arr(1), arr(1)
    progs, _, consts, _ = E.lower(d, -a / 3.0)
    assert tape(progs[0]) == [("push", 0), ("const", 0), ("*", 0), ("const", 1), ("*", 0)]
    assert consts == [-1.0, 1.0 / 3.0]
    a32, d32 = arr(1, torch.float32), arr(1, torch.float32)
    _, _, consts32, _ = E.lower(d32, a32 / 3.0)
    assert consts32 == [float(np.float32(1) / np.float32(3))]


def test_dot_is_right_nested_and_cross_components():
    a, b, c, e = arr(), arr(), arr(), arr(1)
    progs, leaves, consts, _ = E.lower(e, a.cross(b).dot(c) + E.sqrt(a.dot(a) + 1.0))
    t = tape(progs[0])
    # cross comp 2 first (right-nested fold: highest index outermost)
    assert t[:7] == [("push", 0), ("push", 1), ("*", 0), ("push", 2), ("push", 3), ("*", 0), ("-", 0)]
    assert progs[0].n_loads == 9
    assert t.count(("+", 0)) == 2 + 2 + 1 + 1  # two dot folds, +1.0, final sum


def test_random_leaves_are_salted():
    g = ae.Generator.from_seed(42)
    d = arr()
    _, _, _, rngs = E.lower(d, E.uniform(g, 3, -1.0, 1.0))
    assert [r[3] for r in rngs] == [0, 1, 2]
    assert rngs[0][0] == g.key ^ ae.UNIFORM_SALT


def test_limits_are_enforced():
    a, d = arr(1), arr(1)
    x = E.Leaf(a)
    for _ in range(40):
        x = x * 1.0001
    with pytest.raises(ValueError):
        E.lower(d, x)


@pytest.mark.parametrize("dt,mode", [(torch.float64, 0), (torch.float32, 2), (torch.float64, -1)])
def test_tape_kernels_compile_without_a_gpu(dt, mode):
    """The runtime lowering of a tape produces a gfx950 code object (hipRTC needs no device)."""
    import ctypes as C

    from aether_amd import _lib

    a, b, c, d = arr(3, dt), arr(3, dt), arr(3, dt), arr(3, dt)
    g = ae.Generator.from_seed(1)
    x = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)) + E.log(E.abs_(a) + 1.0) / 3.0 - E.uniform(g, 3, 0.0, 1.0)
    progs, _, _, _ = E.lower(d, x)
    size = C.c_uint64()
    _lib.call("aeth_tape_jit_build", 0 if dt == torch.float64 else 1, mode, 3, progs, b"gfx950", C.byref(size))
    assert size.value > 1000


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_wide_tapes_compile_without_a_gpu(dt):
    """Programs with many loads (round-by-round loading) lower and compile too: a 9-plane Vec3 tree and a
    16-plane sum."""
    import ctypes as C

    from aether_amd import _lib

    a, b, c, e = arr(3, dt), arr(3, dt), arr(3, dt), arr(1, dt)
    progs, _, _, _ = E.lower(e, a.cross(b).dot(c) + E.sqrt(a.dot(a) + 1.0))
    assert progs[0].n_loads == 9
    size = C.c_uint64()
    _lib.call("aeth_tape_jit_build", 0 if dt == torch.float64 else 1, 0, 1, progs, b"gfx950", C.byref(size))
    assert size.value > 1000
    many = [arr(1, dt) for _ in range(16)]
    x = many[0]
    for m in many[1:]:
        x = x + m
    progs, _, _, _ = E.lower(many[0], x)
    assert progs[0].n_loads == 16
    _lib.call("aeth_tape_jit_build", 0 if dt == torch.float64 else 1, -1, 1, progs, b"gfx950", C.byref(size))
    assert size.value > 1000
