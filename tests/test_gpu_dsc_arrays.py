"""GPU: aether arrays (aether_amd.dsc) on a HIP device -- the GPU rows of the reference's acceptance tests
(dsc/tests/test_arrays.py): device and pinned memory, fills and copies, torch round trips through DLPack
and __cuda_array_interface__, deleters exactly once, ROCm DLPack stream codes and event ordering, fills
and copies inside a graph capture replayed twice, frees during a capture deferred, stream-ordered frees,
lifetimes at exit. torch stands in for cupy (streams, graphs, a consumer); its busy work stands in for the
reference's spin kernel."""
import ctypes
import gc
import pathlib
import subprocess
import sys
import textwrap
import time

import numpy as np
import pytest
import torch

import aether_amd.dsc as aether
from aether_amd.dsc import _core

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]


def _run(code):
    import os

    env = dict(os.environ, PYTHONPATH=str(ROOT))
    return subprocess.run([sys.executable, "-W", "ignore", "-c", textwrap.dedent(code)], capture_output=True,
                          text=True, timeout=180, env=env)


def host(a):
    return np.array(aether.asarray(a, device="cpu"), copy=True)


def busy(stream, ms_target=600):
    """Queue roughly `ms_target` ms of matrix products on `stream` (no memory of the arrays under test)."""
    with torch.cuda.stream(stream):
        x = torch.randn(4096, 4096, device="cuda")
        y = x
        for _ in range(max(1, ms_target // 2)):
            y = x @ x
    return y


def test_gpu_surface_copy_and_fill(dev):
    d = aether.zeros((3, 5), "float32", "cuda")
    assert d.device == aether.Device("cuda", 0) and d.nbytes == 60 and d.strides == (20, 4)
    assert not host(d).any()
    for dt, val in (("float64", 0.1), ("float32", 2.5), ("int64", -3), ("int32", 7), ("uint64", 5), ("uint32", 6),
                    ("uint8", 200), ("bool", True), ("int8", -5), ("int16", -300), ("uint16", 60000),
                    ("float16", 1.5)):
        f = aether.full(33, val, dt, "cuda")
        got = host(f)
        assert got.dtype == np.dtype(dt) and np.all(got == np.array(val, dt)), (dt, val)
    bf = aether.full(5, 1.0, "bfloat16", "cuda")
    back = aether.empty(5, "bfloat16")
    aether.copy(bf, into=back)
    assert ctypes.string_at(back.ptr, back.nbytes) == b"\x80\x3f" * 5
    h = aether.pinned((3, 5), "float32")
    aether.copy(aether.full((3, 5), 2.5, "float32", "cuda"), into=h)
    assert set(memoryview(h).tolist()[0]) == {2.5}
    up = aether.asarray(aether.asarray([[1.0, 2.0], [3.0, 4.0]]), device="cuda")
    assert aether.copy(up, into=aether.empty((2, 2), device="cpu")).shape == (2, 2)
    assert host(up).tolist() == [[1.0, 2.0], [3.0, 4.0]]
    assert aether.asarray(up) is up and aether.zeros_like(up).device == up.device
    assert aether.cuda_plugin() == "hip" and aether.cuda_runtime_version() > 0
    with pytest.raises(aether.Error):
        aether.copy(up, into=aether.zeros(5))
    dd = aether.copy(up)  # device to device
    assert dd.device == up.device and host(dd).tolist() == [[1.0, 2.0], [3.0, 4.0]]


def test_device_memory_has_no_buffer(dev):
    with pytest.raises(BufferError):
        memoryview(aether.zeros(4, device="cuda"))


def test_pinned_is_host_memory_zero_copy(dev):
    a = aether.zeros((3, 7), "float64", "pinned")
    assert a.device == aether.Device("pinned", 0) and a.__dlpack_device__() == (11, 0)
    m = memoryview(a)
    m[1, 2] = 4.0
    v = np.asarray(a)
    assert v.ctypes.data == a.ptr and v[1, 2] == 4.0
    v[0, 0] = 9.0
    assert memoryview(a)[0, 0] == 9.0
    a.fill(3.0)
    assert float(v.sum()) == 63.0


def test_torch_roundtrip_shares_memory(dev):
    a = aether.full(8, 3.0, "float32", "cuda")
    a.synchronize()
    assert a.__dlpack_device__() == (10, 0)
    t = torch.from_dlpack(a)
    assert t.data_ptr() == a.ptr and t.device.type == "cuda" and float(t.sum()) == 24.0
    t *= 2.0  # a torch kernel writes the shared memory
    torch.cuda.synchronize()
    assert host(a).tolist() == [6.0] * 8
    src = torch.arange(12, dtype=torch.float64, device="cuda").reshape(3, 4)
    b = aether.from_dlpack(src)
    assert b.shape == (3, 4) and b.ptr == src.data_ptr() and host(b).tolist() == src.cpu().tolist()
    col = aether.from_dlpack(src[:, 1])  # a strided view keeps its strides
    assert col.shape == (3,) and col.strides == (32,) and not col.is_contiguous
    s = torch.cuda.Stream()
    c = aether.from_dlpack(torch.ones(4, device="cuda"), stream=s)
    assert c.shape == (4,)


def test_cuda_array_interface(dev):
    a = aether.zeros((2, 3), "int32", "cuda")
    cai = a.__cuda_array_interface__
    assert cai["shape"] == (2, 3) and cai["typestr"] == "<i4" and cai["data"] == (a.ptr, False)
    assert cai["version"] == 3 and cai["strides"] is None and cai["stream"] is None
    assert not hasattr(aether.zeros(2), "__cuda_array_interface__")
    t = torch.arange(6, dtype=torch.int32, device="cuda").reshape(2, 3)
    v = aether.asarray(t)  # torch exports __cuda_array_interface__: viewed without a copy
    assert v.ptr == t.data_ptr() and v.shape == (2, 3) and host(v).tolist() == [[0, 1, 2], [3, 4, 5]]
    tt = t.t()
    w = aether.asarray(tt)
    assert w.strides == (4, 12) and not w.is_contiguous


def test_fill_of_an_unaligned_device_view(dev):
    base = torch.zeros(40, dtype=torch.float64, device="cuda")
    for lo, hi in ((1, 38), (3, 4), (1, 2), (0, 40), (5, 40)):
        base.zero_()
        v = aether.asarray(base[lo:hi])
        v.fill(2.5)
        torch.cuda.synchronize()
        want = torch.zeros(40, dtype=torch.float64)
        want[lo:hi] = 2.5
        assert torch.equal(base.cpu(), want), (lo, hi)
    bytes_ = torch.zeros(64, dtype=torch.uint8, device="cuda")
    v = aether.asarray(bytes_[3:50])
    v.fill(7)
    torch.cuda.synchronize()
    want = torch.zeros(64, dtype=torch.uint8)
    want[3:50] = 7
    assert torch.equal(bytes_.cpu(), want)


def test_deleter_once_both_generations(dev):
    before = _core._exports_live()
    a = aether.zeros(16, "float32", "cuda")
    cap_v, cap_l = a.__dlpack__(max_version=(1, 0)), a.__dlpack__()
    assert _core._exports_live() == before + 2
    del cap_v, cap_l  # never consumed: the capsule destructor releases
    gc.collect()
    assert _core._exports_live() == before
    view = torch.from_dlpack(a)  # consumed: the consumer's deleter releases
    assert _core._exports_live() == before + 1
    del a
    view += 1.0  # the export keeps the memory alive
    torch.cuda.synchronize()
    assert float(view.sum()) == 16.0
    del view
    gc.collect()
    assert _core._exports_live() == before


def test_import_deleter_fires_once(dev):
    c = torch.zeros(8, device="cuda")
    a = aether.from_dlpack(c)
    del a
    gc.collect()
    assert float(c.sum()) == 0.0  # the producer's memory is intact and still its own


def test_dlpack_stream_codes_and_event_ordering(dev):
    a = aether.zeros(8, "float64", "cuda")
    for bad in (1, 2):  # not ROCm stream codes
        with pytest.raises(ValueError):
            a.__dlpack__(stream=bad)
    for code in (None, -1, 0):
        a.__dlpack__(stream=code)
    own, prod = torch.cuda.Stream(), torch.cuda.Stream()
    b = aether.zeros(8, "float64", "cuda", stream=prod)
    tb = torch.from_dlpack(b)
    keep = busy(prod)
    with torch.cuda.stream(prod):
        tb.fill_(1.0)  # after the busy work, on the producer stream
    t = time.perf_counter()
    b.__dlpack__(stream=own.cuda_stream)  # the consumer stream waits on an event; the host does not
    waited = time.perf_counter() - t
    with torch.cuda.stream(own):
        got = torch.from_dlpack(b).clone()  # ordered after the producer's fill
    own.synchronize()
    assert waited < 0.2, f"the host waited {waited:.3f}s for the producer's work"
    assert float(got.sum()) == 8.0
    del keep
    prod.synchronize()


def test_zero_fill_capture_two_replays(dev):
    s = torch.cuda.Stream()
    d = aether.empty(4096, "float64", "cuda", stream=s)
    w = aether.empty(4096, "float64", "cuda", stream=s)
    pin = aether.pinned(4096, "float64")
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        d.zero(s)
        w.fill(0.1, s)
        aether.copy(d, into=pin, stream=s)
    td, tw = torch.from_dlpack(d), torch.from_dlpack(w)
    for _ in range(2):
        td.fill_(7.0)
        tw.fill_(3.0)
        pin.fill(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert not any(memoryview(pin).tolist())
        assert float(tw.min()) == 0.1 == float(tw.max())


@pytest.mark.parametrize("own_stream", [False, True])
def test_gc_free_inside_global_capture(dev, own_stream):
    out = _run(f"""
        import gc, torch
        import aether_amd.dsc as aether
        own = torch.cuda.Stream()
        a = aether.zeros(1024, "float64", "cuda", stream={'own' if own_stream else 'None'})
        cyc = [a]; cyc.append(cyc)
        del a, cyc
        torch.cuda.synchronize()
        gc.disable()
        x = torch.zeros(8, device="cuda")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="global"):
            x += 1
            gc.collect()  # frees the array while the capture runs
            x += 1
        g.replay()
        torch.cuda.synchronize()
        assert float(x.sum()) == 16.0, float(x.sum())
        aether.empty(1, device="cuda")  # the next allocation outside the capture drains the deferred free
        print("ok")
    """)
    assert out.returncode == 0 and out.stdout.strip() == "ok", (out.returncode, out.stdout, out.stderr[-800:])


@pytest.mark.parametrize("mode", ["legacy", "own_stream"])
def test_early_drop_after_launch_is_stream_ordered(dev, mode):
    own = torch.cuda.Stream()
    stream = own if mode == "own_stream" else torch.cuda.default_stream()
    a = aether.zeros(8, "float64", "cuda", stream=own if mode == "own_stream" else None)
    keep = busy(stream)
    with torch.cuda.stream(stream):
        t = torch.from_dlpack(a)
        t.fill_(7.0)
        del t
    t0 = time.perf_counter()
    del a
    gc.collect()
    waited = time.perf_counter() - t0
    stream.synchronize()
    del keep
    assert waited < 0.5, f"the free blocked the host for {waited:.2f}s: it was not queued on the stream"


def test_arrays_alive_at_exit_and_stream_object_held(dev):
    out = _run("""
        import torch
        import aether_amd.dsc as aether
        s = torch.cuda.Stream()
        a = aether.zeros(1024, "float64", "cuda", stream=s)
        b = aether.zeros(1024, "float64", "cuda")
        p = aether.pinned(8)
        del s  # the array still holds the stream object
        e = aether.zeros(512, "float32", "cuda", stream=torch.cuda.Stream())
        del e
        print("ok")
    """)
    assert out.returncode == 0 and out.stdout.strip() == "ok", (out.returncode, out.stderr[-600:])


def test_alloc_free_timing_reported(dev, capsys):
    n = 1000
    t = time.perf_counter()
    xs = [aether.empty(64, device="cuda") for _ in range(n)]
    ta = (time.perf_counter() - t) / n
    t = time.perf_counter()
    del xs
    tf = (time.perf_counter() - t) / n
    t = time.perf_counter()
    ys = [torch.empty(64, device="cuda") for _ in range(n)]
    tt = (time.perf_counter() - t) / n
    del ys
    with capsys.disabled():
        print(f"\naether alloc {ta * 1e6:.1f} us  free {tf * 1e6:.1f} us  torch.empty {tt * 1e6:.1f} us")
