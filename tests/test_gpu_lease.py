"""GPU: the per-stream scratch lease that the fused sums' partials take outside a stream capture.

A call's partials live in one growable buffer per (device, stream), reused by every later call on that
stream. The results must not depend on it: calls of different sizes back to back on one stream, calls from
two host threads into one stream, and calls on several streams at once all give the bits the same call
gives alone on a fresh stream. Issuing a call must also leave the host free while the GPU works.
"""
import ctypes as C
import threading
import time

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def _inputs(n, dt=torch.float64):
    arrs = []
    for s in (11, 12, 13):
        a = ae.Array(n, 3, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        arrs.append(a)
    return arrs


def _values(d):
    """The n samples of each plane (the capacity padding past n is never written)."""
    return d.data[:, :d.n].cpu().numpy().copy()


def _tree(a, b, c):
    return a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))


def _alone(n, mode=ae.SUM_NEUMAIER, dt=torch.float64):
    """Values and sums of the tape evaluation and of the specialised kernel, each on a fresh stream."""
    a, b, c = _inputs(n, dt)
    torch.cuda.synchronize()
    out = []
    for path in ("tape", "cfg0"):
        d = ae.Array(n, 3, dt)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            if path == "tape":
                sums = E.assign(d, _tree(a, b, c), mode)
            else:
                sums = ae.eval_cfg0(a, b, c, d, mode)
        s.synchronize()
        out.append((_values(d), sums.cpu().numpy().copy()))
    return (a, b, c), out


SIZES = [100_003, 1 << 22, 5, 3 * (1 << 20) + 17, 2049]


def test_sizes_back_to_back_on_one_stream(dev):
    cases = {n: _alone(n) for n in SIZES}
    torch.cuda.synchronize()
    # grow and shrink the stream's buffer: every call queued before any result is read
    pending = []
    for n in SIZES + SIZES[::-1]:
        (a, b, c), _ = cases[n]
        d1, d2 = ae.Array(n, 3, torch.float64), ae.Array(n, 3, torch.float64)
        s1 = E.assign(d1, _tree(a, b, c), ae.SUM_NEUMAIER)
        s2 = ae.eval_cfg0(a, b, c, d2, ae.SUM_NEUMAIER)
        pending.append((n, d1, s1, d2, s2))
    torch.cuda.synchronize()
    for n, d1, s1, d2, s2 in pending:
        (_, _, _), [(v1, w1), (v2, w2)] = cases[n]
        assert np.array_equal(_values(d1), v1) and np.array_equal(s1.cpu().numpy(), w1), n
        assert np.array_equal(_values(d2), v2) and np.array_equal(s2.cpu().numpy(), w2), n


def test_two_host_threads_one_stream(dev):
    n = 1 << 21
    (a, b, c), [(v1, w1), _] = _alone(n)
    stream = torch.cuda.Stream()
    errors, outs = [], [[], []]

    def worker(k):
        try:
            with torch.cuda.stream(stream):
                for _ in range(12):
                    d = ae.Array(n, 3, torch.float64)
                    outs[k].append((d, E.assign(d, _tree(a, b, c), ae.SUM_NEUMAIER)))
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    stream.synchronize()
    for k in range(2):
        for d, s in outs[k]:
            assert np.array_equal(s.cpu().numpy(), w1)
            assert np.array_equal(_values(d), v1)


def test_streams_at_once(dev):
    n = (1 << 20) + 5
    (a, b, c), [(v1, w1), (v2, w2)] = _alone(n)
    streams = [torch.cuda.Stream() for _ in range(4)]
    res = []
    for rep in range(3):
        for s in streams:
            with torch.cuda.stream(s):
                d = ae.Array(n, 3, torch.float64)
                e = ae.Array(n, 3, torch.float64)
                res.append((d, E.assign(d, _tree(a, b, c), ae.SUM_NEUMAIER), e,
                            ae.eval_cfg0(a, b, c, e, ae.SUM_NEUMAIER)))
    torch.cuda.synchronize()
    for d, s1, e, s2 in res:
        assert np.array_equal(s1.cpu().numpy(), w1) and np.array_equal(_values(d), v1)
        assert np.array_equal(s2.cpu().numpy(), w2) and np.array_equal(_values(e), v2)


def test_float_float_sums_share_the_lease(dev):
    n = 300_001
    (a, b, c), [(v1, w1), (v2, w2)] = _alone(n, ae.SUM_FLOATFLOAT, torch.float32)
    for _ in range(3):
        d, e = ae.Array(n, 3, torch.float32), ae.Array(n, 3, torch.float32)
        s1 = E.assign(d, _tree(a, b, c), ae.SUM_FLOATFLOAT)
        s2 = ae.eval_cfg0(a, b, c, e, ae.SUM_FLOATFLOAT)
        torch.cuda.synchronize()
        assert np.array_equal(s1.cpu().numpy(), w1) and np.array_equal(_values(d), v1)
        assert np.array_equal(s2.cpu().numpy(), w2) and np.array_equal(_values(e), v2)


def test_host_is_not_held(dev):
    """Queuing evaluations with sums returns long before the GPU has run them."""
    n = 1 << 26
    a, b, c = _inputs(n)
    d = ae.Array(n, 3, torch.float64)
    sums = torch.empty((3, 2), dtype=torch.float64, device="cuda")
    tree = _tree(a, b, c)
    E.assign(d, tree, ae.SUM_NEUMAIER, sums)
    ae.eval_cfg0(a, b, c, d, ae.SUM_NEUMAIER, sums)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    t0 = time.perf_counter()
    for _ in range(4):
        E.assign(d, tree, ae.SUM_NEUMAIER, sums)
        ae.eval_cfg0(a, b, c, d, ae.SUM_NEUMAIER, sums)
    host_ms = (time.perf_counter() - t0) * 1e3
    e1.record()
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1)
    assert host_ms < 0.5 * gpu_ms, (host_ms, gpu_ms)


def test_idle_streams_release_their_buffers(dev):
    """48 streams of the HIP runtime's own (distinct handles), one call each: the device's default memory
    pool, where the buffers live, stays near its use after the first 20 instead of growing by a buffer a
    stream, and every call gives the bits of the call alone."""
    hip = C.CDLL("libamdhip64.so")
    n = 1 << 22
    (a, b, c), [(v1, w1), _] = _alone(n)
    tree = _tree(a, b, c)
    pool = C.c_void_p()
    assert hip.hipDeviceGetDefaultMemPool(C.byref(pool), 0) == 0

    def used():
        v = C.c_uint64()
        assert hip.hipMemPoolGetAttribute(pool, 0x7, C.byref(v)) == 0  # hipMemPoolAttrUsedMemCurrent
        return v.value

    streams, base = [], None
    try:
        for k in range(48):
            s = C.c_void_p()
            assert hip.hipStreamCreate(C.byref(s)) == 0
            streams.append(s)
            ext = torch.cuda.ExternalStream(s.value)
            d = ae.Array(n, 3, torch.float64)
            with torch.cuda.stream(ext):
                sums = E.assign(d, tree, ae.SUM_NEUMAIER)
            ext.synchronize()
            assert np.array_equal(sums.cpu().numpy(), w1) and np.array_equal(_values(d), v1), k
            if k == 19:
                base = used()
        torch.cuda.synchronize()
        # one buffer a stream is 1 MiB here: 28 more streams unreleased would add 28 MiB
        assert used() <= base + (4 << 20), (used(), base)
    finally:
        torch.cuda.synchronize()
        for s in streams:
            hip.hipStreamDestroy(s)
