"""GPU: the per-launch timing of the tape kernels (aeth_tape_timing / aeth_tape_timing_read).

One pair of HIP events per tape kernel launched while timing is on, none while it is off or inside a stream
capture; reads hand back the oldest launches first and leave the rest pending; the timed kernel takes no
longer than the whole call around it, and timing changes no bit of the results."""
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def _inputs(n):
    out = []
    for s in (31, 32, 33):
        a = ae.Array(n, 3, torch.float64)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        out.append(a)
    return out


def _tree(a, b, c):
    return a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))


def test_one_pair_per_launch_and_same_bits(dev):
    n = 1 << 22
    a, b, c = _inputs(n)
    tree = _tree(a, b, c)
    ref = ae.Array(n, 3, torch.float64)
    want = E.assign(ref, tree, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    E.kernel_times()
    E.kernel_timing(True)
    try:
        outs = []
        for _ in range(5):
            d = ae.Array(n, 3, torch.float64)
            outs.append((d, E.assign(d, tree, ae.SUM_NEUMAIER)))
        E.assign(ae.Array(n, 3, torch.float64), tree)  # no sums: still one tape launch
    finally:
        E.kernel_timing(False)
    E.assign(ae.Array(n, 3, torch.float64), tree, ae.SUM_NEUMAIER)  # after timing is off: not recorded
    ms = E.kernel_times()
    assert len(ms) == 6
    assert all(t > 0.0 for t in ms)
    assert E.kernel_times() == []
    for d, s in outs:
        assert torch.equal(s, want)
        assert torch.equal(d.data[:, :n], ref.data[:, :n])


def test_partial_reads_keep_order(dev):
    n = 1 << 20
    a, b, c = _inputs(n)
    d = ae.Array(n, 3, torch.float64)
    E.kernel_times()
    E.kernel_timing(True)
    try:
        for _ in range(4):
            E.assign(d, _tree(a, b, c), ae.SUM_NEUMAIER)
    finally:
        E.kernel_timing(False)
    import ctypes as C
    count = C.c_int32()
    buf = (C.c_float * 3)()
    ae._lib.call("aeth_tape_timing_read", buf, 3, C.byref(count))
    assert count.value == 4
    assert all(buf[k] > 0.0 for k in range(3))
    rest = E.kernel_times()
    assert len(rest) == 1 and rest[0] > 0.0


def test_kernel_within_the_call(dev):
    """The tape kernel's events sit inside events around the whole call (the sums' tree is after it)."""
    n = 1 << 25
    a, b, c = _inputs(n)
    d = ae.Array(n, 3, torch.float64)
    sums = torch.empty((3, 2), dtype=torch.float64, device="cuda")
    tree = _tree(a, b, c)
    for _ in range(3):
        E.assign(d, tree, ae.SUM_NEUMAIER, sums)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(6)]
    E.kernel_times()
    E.kernel_timing(True)
    try:
        for e0, e1 in ev:
            e0.record()
            E.assign(d, tree, ae.SUM_NEUMAIER, sums)
            e1.record()
    finally:
        E.kernel_timing(False)
    torch.cuda.synchronize()
    ms = E.kernel_times()
    call_ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    assert len(ms) == len(call_ms)
    for k, c_ms in zip(ms, call_ms):
        assert 0.0 < k <= c_ms * 1.001 + 0.005, (k, c_ms)


def test_not_recorded_under_capture(dev):
    n = 1 << 18
    a, b, c = _inputs(n)
    d = ae.Array(n, 3, torch.float64)
    tree = _tree(a, b, c)
    E.prepare(d, tree, ae.SUM_NEUMAIER)
    E.assign(d, tree, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    want = d.data[:, :n].clone()
    E.kernel_times()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    E.kernel_timing(True)
    try:
        with torch.cuda.graph(g, stream=s):
            E.assign(d, tree, ae.SUM_NEUMAIER)
    finally:
        E.kernel_timing(False)
    assert E.kernel_times() == []
    d.data.zero_()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(d.data[:, :n], want)
