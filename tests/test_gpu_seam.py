"""GPU: the `aether-memory/1` seam (include/aether_hip.h, mirroring the
reference's dsc/src/seam.h) behaves as the reference's plugin does
(dsc/src/plugin.cpp): stream-ordered chunks, fill with a pattern, ordered
copies between kinds, deferred frees under capture, FFT plans over axes
(vs the long-double DFT oracle with the reference's bound, and numpy.fft).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from aether_amd import _lib

pytestmark = pytest.mark.gpu

CPU, DEVICE, PINNED = 0, 1, 2
_vp, _i, _sz = C.c_void_p, C.c_int, C.c_size_t


class Seam(C.Structure):
    _fields_ = [
        ("abi", C.c_uint32),
        ("probe", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("runtime_version", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("current_device", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("last_error", C.CFUNCTYPE(C.c_char_p)),
        ("chunk_alloc", C.CFUNCTYPE(_i, _i, _i, _sz, _vp, C.POINTER(_vp), C.POINTER(_vp))),
        ("chunk_release", C.CFUNCTYPE(None, _vp)),
        ("drain", C.CFUNCTYPE(None)),
        ("fill", C.CFUNCTYPE(_i, _i, _i, _vp, _sz, _vp, _sz, _vp)),
        ("copy", C.CFUNCTYPE(_i, _i, _i, _vp, _i, _i, _vp, _sz, _vp)),
        ("stream_wait", C.CFUNCTYPE(_i, _i, _vp, _vp)),
        ("stream_sync", C.CFUNCTYPE(_i, _i, _vp)),
        ("pointer_device", C.CFUNCTYPE(_i, _vp, C.POINTER(_i))),
        ("fft_plan", C.CFUNCTYPE(_i, _i, _i, _i, _i, _i, _i, C.POINTER(C.c_int64), _i, C.POINTER(C.c_int32), _vp,
                                 C.POINTER(_vp))),
        ("fft_exec", C.CFUNCTYPE(_i, _vp, _vp, _vp, _vp)),
        ("fft_destroy", C.CFUNCTYPE(None, _vp)),
    ]


@pytest.fixture(scope="module")
def seam(dev):
    L = _lib.lib()
    fn = L.aether_memory_v1
    fn.restype = C.POINTER(Seam)
    t = fn().contents
    assert t.abi == 2
    n = _i()
    assert t.probe(C.byref(n)) == 0 and n.value >= 1
    return t


def alloc(seam, kind, nbytes, stream=None):
    ch, p = _vp(), _vp()
    assert seam.chunk_alloc(kind, 0, nbytes, stream, C.byref(ch), C.byref(p)) == 0, seam.last_error()
    return ch, p.value


def test_fill_copy_roundtrip(seam):
    n = 1 << 20
    s = torch.cuda.current_stream().cuda_stream
    dch, d = alloc(seam, DEVICE, n * 8, s)
    hch, h = alloc(seam, PINNED, n * 8)
    val = C.c_double(3.25)
    assert seam.fill(DEVICE, 0, d, n * 8, C.byref(val), 8, s) == 0
    assert seam.copy(PINNED, 0, h, DEVICE, 0, d, n * 8, s) == 0
    assert seam.stream_sync(0, s) == 0
    arr = np.ctypeslib.as_array(C.cast(h, C.POINTER(C.c_double)), shape=(n,))
    assert np.all(arr == 3.25)
    # host -> device -> device -> host with a pattern, unaligned tail
    src = np.arange(n, dtype=np.float64)
    d2ch, d2 = alloc(seam, DEVICE, n * 8, s)
    assert seam.copy(DEVICE, 0, d, CPU, 0, src.ctypes.data, n * 8 - 8, s) == 0
    assert seam.copy(DEVICE, 0, d2, DEVICE, 0, d, n * 8, s) == 0  # carries d's filled last element
    out = np.zeros(n)
    assert seam.copy(CPU, 0, out.ctypes.data, DEVICE, 0, d2, n * 8, s) == 0
    assert np.array_equal(out[:-1], src[:-1]) and out[-1] == 3.25
    dv = _i(-1)
    assert seam.pointer_device(d, C.byref(dv)) == 0 and dv.value == 0
    for ch in (dch, hch, d2ch):
        seam.chunk_release(ch)


def test_fill_patterns_and_errors(seam):
    s = torch.cuda.current_stream().cuda_stream
    ch, d = alloc(seam, DEVICE, 4096, s)
    pat = (C.c_uint8 * 4)(1, 2, 3, 4)
    assert seam.fill(DEVICE, 0, d, 4096, pat, 4, s) == 0
    out = np.zeros(4096, np.uint8)
    assert seam.copy(CPU, 0, out.ctypes.data, DEVICE, 0, d, 4096, s) == 0
    assert np.array_equal(out, np.tile([1, 2, 3, 4], 1024))
    assert seam.fill(DEVICE, 0, d, 4095, pat, 4, s) == 1  # byte count not a multiple of the width
    assert b"width" in seam.last_error()
    seam.chunk_release(ch)


# Byte counts around the bulk kernels' 16 KiB units: below one unit, whole units, units plus spare vectors
# plus tail bytes, and regions past the streaming threshold (fill 512 MiB, copy 256 MiB).
BULK_SIZES = [1, 15, 16, 4095, 16384, 16384 + 16, 3 * 16384 + 1008 + 7, (1 << 22) + 48, (520 << 20) + 5]


@pytest.mark.parametrize("nbytes", BULK_SIZES)
@pytest.mark.parametrize("offset", [0, 3, 8])
def test_bulk_fill_and_copy_bytes(seam, nbytes, offset):
    """Device fill (1-byte and 8-byte patterns) and same-device copy against torch's own result, byte for
    byte; `offset` moves the start off the 16-byte grid (the fill's bytewise head, the copy's byte path)."""
    s = torch.cuda.current_stream().cuda_stream
    buf = torch.zeros(nbytes + 32, dtype=torch.uint8, device="cuda")
    base = buf.data_ptr() + offset
    b1 = C.c_uint8(0xA5)
    assert seam.fill(DEVICE, 0, base, nbytes, C.byref(b1), 1, s) == 0, seam.last_error()
    torch.cuda.synchronize()
    assert torch.all(buf[offset:offset + nbytes] == 0xA5)
    assert torch.all(buf[:offset] == 0) and torch.all(buf[offset + nbytes:] == 0)
    if offset % 8 == 0 and nbytes % 8 == 0:
        pat = (C.c_uint8 * 8)(*range(1, 9))
        assert seam.fill(DEVICE, 0, base, nbytes, pat, 8, s) == 0, seam.last_error()
        torch.cuda.synchronize()
        want = torch.arange(1, 9, dtype=torch.uint8, device="cuda").repeat(nbytes // 8)
        assert torch.equal(buf[offset:offset + nbytes], want)
    src = torch.randint(0, 256, (nbytes + 32,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros_like(src)
    assert seam.copy(DEVICE, 0, dst.data_ptr() + offset, DEVICE, 0, src.data_ptr() + offset, nbytes, s) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst[offset:offset + nbytes], src[offset:offset + nbytes])
    assert torch.all(dst[:offset] == 0) and torch.all(dst[offset + nbytes:] == 0)
    del buf, src, dst


def test_fill_and_copy_in_graph_replays(seam):
    # capture-legal: replay several times with eager work between, every replay checked
    n = 1 << 16
    a = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    g = torch.cuda.CUDAGraph()
    v = C.c_double(0.0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        s = torch.cuda.current_stream().cuda_stream
        v.value = 7.5
        assert seam.fill(DEVICE, 0, a.data_ptr(), n * 8, C.byref(v), 8, s) == 0
        assert seam.copy(DEVICE, 0, b.data_ptr(), DEVICE, 0, a.data_ptr(), n * 8, s) == 0
    for k in range(4):
        a.fill_(float(k))
        b.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.all(a == 7.5) and torch.all(b == 7.5)


def test_release_under_capture_is_deferred(seam):
    side = torch.cuda.Stream()
    ch, _ = alloc(seam, DEVICE, 1 << 20, side.cuda_stream)
    x = torch.zeros(16, device="cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        x.add_(1.0)
        seam.chunk_release(ch)  # queued, not freed inside the capture
    seam.drain()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert torch.all(x == 1.0)


def test_fft_destroy_under_capture(seam):
    """fft_destroy is safe during a capture (dsc/src/seam.h:48-50): destroying one plan while another
    plan's execution is being captured (global capture mode, torch's default) neither breaks the
    capture nor the replays; the destroyed plan's buffers are freed once the capture has ended."""
    shp = (C.c_int64 * 2)(8, 1024)
    ax = (C.c_int32 * 1)(1)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((8, 1024)) + 1j * rng.standard_normal((8, 1024))
    xin = torch.from_numpy(x).to("cuda")
    out = torch.zeros_like(xin)
    side = torch.cuda.Stream()
    s = side.cuda_stream
    keep, doomed, doomed_real = _vp(), _vp(), _vp()
    assert seam.fft_plan(0, 0, 1, 0, 1, 2, shp, 1, ax, s, C.byref(keep)) == 0, seam.last_error()
    assert seam.fft_plan(0, 0, 1, 0, 1, 2, shp, 1, ax, s, C.byref(doomed)) == 0, seam.last_error()
    big = (C.c_int64 * 1)(1 << 14)  # four-step plan: work arrays and twiddle tables of its own
    ax0 = (C.c_int32 * 1)(0)
    assert seam.fft_plan(0, 1, 1, 0, 1, 1, big, 1, ax0, s, C.byref(doomed_real)) == 0, seam.last_error()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        assert seam.fft_exec(keep, xin.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream) == 0
        seam.fft_destroy(doomed)
        seam.fft_destroy(doomed_real)
    seam.drain()
    ref = np.fft.fft(x, axis=1)
    for _ in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-10 * np.abs(ref).max()
    seam.fft_destroy(keep)
    seam.drain()


def test_batch_plan_destroy_under_capture(dev):
    L = _lib.lib()
    lengths = (C.c_uint64 * 2)(64, 128)
    offs = (C.c_uint64 * 2)(0, 64)
    keep, doomed = _vp(), _vp()
    assert L.aeth_fft_batch_plan(0, 2, lengths, offs, offs, C.byref(keep)) == 0
    assert L.aeth_fft_batch_plan(0, 2, lengths, offs, offs, C.byref(doomed)) == 0
    rng = np.random.default_rng(12)
    x = rng.standard_normal(192) + 1j * rng.standard_normal(192)
    xin = torch.from_numpy(x).to("cuda")
    out = torch.zeros_like(xin)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        st = torch.cuda.current_stream().cuda_stream
        assert L.aeth_fft_batch_exec(keep, _vp(xin.data_ptr()), _vp(out.data_ptr()), 1, 0, _vp(st)) == 0
        assert L.aeth_fft_batch_destroy(doomed) == 0
    ref = np.concatenate([np.fft.fft(x[:64]), np.fft.fft(x[64:])])
    for _ in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-12 * np.abs(ref).max()
    assert L.aeth_fft_batch_destroy(keep) == 0


@pytest.mark.parametrize("shape,axes", [((8, 64), (1,)), ((64, 8), (0,)), ((16, 32), (0, 1)), ((4, 8, 16), (1,)),
                                        ((4, 8, 16), (0, 2)), ((1024,), (0,))])
@pytest.mark.parametrize("norm,forward", [(0, 1), (0, 0), (2, 1)])
def test_fft_plan_axes(seam, orc, shape, axes, norm, forward):
    rng = np.random.default_rng(len(shape) * 10 + len(axes))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    xin = torch.from_numpy(x).to("cuda")
    out = torch.zeros_like(xin)
    s = torch.cuda.current_stream().cuda_stream
    shp = (C.c_int64 * len(shape))(*shape)
    ax = (C.c_int32 * len(axes))(*axes)
    plan = _vp()
    assert seam.fft_plan(0, 0, 1, norm, forward, len(shape), shp, len(axes), ax, s, C.byref(plan)) == 0, \
        seam.last_error()
    assert seam.fft_exec(plan, xin.data_ptr(), out.data_ptr(), s) == 0
    torch.cuda.synchronize()
    seam.fft_destroy(plan)
    y = out.cpu().numpy()
    npn = {0: "backward", 1: "forward", 2: "ortho"}[norm]
    ref = np.fft.fftn(x, axes=axes, norm=npn) if forward else np.fft.ifftn(x, axes=axes, norm=npn)
    points = int(np.prod([shape[a] for a in axes]))
    bound = (8 + 4 * np.log2(points + 1)) * 2.0**-52 * np.linalg.norm(x) * max(1.0, np.abs(ref).max() /
                                                                                max(np.abs(x).max(), 1e-300))
    assert np.max(np.abs(y - ref)) <= bound
    if len(shape) == 1 and forward and norm == 0:
        assert np.max(np.abs(y - orc.dft(x))) <= (8 + 4 * np.log2(points + 1)) * 2.0**-52 * np.linalg.norm(x)


def test_fft_plan_unsupported(seam):
    s = torch.cuda.current_stream().cuda_stream
    shp = (C.c_int64 * 1)((1 << 32) + 1)  # not a power of two, and not below Bluestein's 2^32
    ax = (C.c_int32 * 1)(0)
    plan = _vp()
    assert seam.fft_plan(0, 0, 1, 0, 1, 1, shp, 1, ax, s, C.byref(plan)) == 2
    assert seam.fft_plan(0, 1, 1, 0, 1, 1, shp, 1, ax, s, C.byref(plan)) == 2  # R2C of that length
    shp[0] = 1 << 35  # a power of two beyond the four-step's 2^34
    assert seam.fft_plan(0, 0, 1, 0, 1, 1, shp, 1, ax, s, C.byref(plan)) == 2
    assert seam.fft_plan(0, 3, 1, 0, 1, 1, shp, 1, ax, s, C.byref(plan)) == 1  # unknown kind
    ax2 = (C.c_int32 * 2)(0, 0)
    shp2 = (C.c_int64 * 2)(4, 4)
    assert seam.fft_plan(0, 0, 1, 0, 1, 2, shp2, 2, ax2, s, C.byref(plan)) == 1  # repeated axis


@pytest.mark.parametrize("shape,axes", [((8, 64), (1,)), ((64, 8), (0,)), ((16, 32), (0, 1)), ((16, 32), (1, 0)),
                                        ((4, 8, 16), (1,)), ((4, 8, 16), (0, 2)), ((2048,), (0,)), ((3, 2), (1,)),
                                        ((5, 1), (1,))])
@pytest.mark.parametrize("norm", [0, 1, 2])
@pytest.mark.parametrize("dbl", [1, 0])
def test_fft_plan_real(seam, shape, axes, norm, dbl):
    """R2C then C2R vs numpy's rfftn / irfftn (cufft_plan.h:9-12 layout; fft/Fft.h:198-236)."""
    rng = np.random.default_rng(sum(shape) + len(axes))
    x = rng.standard_normal(shape)
    half = list(shape)
    half[axes[-1]] = shape[axes[-1]] // 2 + 1
    ft = np.float64 if dbl else np.float32
    ct = np.complex128 if dbl else np.complex64
    xin = torch.from_numpy(x.astype(ft)).to("cuda")
    spec = torch.zeros(tuple(half), dtype=torch.complex128 if dbl else torch.complex64, device="cuda")
    back = torch.zeros_like(xin)
    s = torch.cuda.current_stream().cuda_stream
    shp = (C.c_int64 * len(shape))(*shape)
    ax = (C.c_int32 * len(axes))(*axes)
    npn = {0: "backward", 1: "forward", 2: "ortho"}[norm]
    p1, p2 = _vp(), _vp()
    assert seam.fft_plan(0, 1, dbl, norm, 0, len(shape), shp, len(axes), ax, s, C.byref(p1)) == 0, seam.last_error()
    assert seam.fft_plan(0, 2, dbl, norm, 1, len(shape), shp, len(axes), ax, s, C.byref(p2)) == 0, seam.last_error()
    assert seam.fft_exec(p1, xin.data_ptr(), spec.data_ptr(), s) == 0, seam.last_error()
    spec_in = spec.clone()
    assert seam.fft_exec(p2, spec.data_ptr(), back.data_ptr(), s) == 0, seam.last_error()
    torch.cuda.synchronize()
    seam.fft_destroy(p1)
    seam.fft_destroy(p2)
    eps = 2.0**-52 if dbl else 2.0**-23
    points = int(np.prod([shape[a] for a in axes]))
    xr = x.astype(ft).astype(np.float64)
    # a stable FFT: ||dy||_2 <= c log2(N) eps ||y||_2 with ||y||_2 = sqrt(N) s ||x||_2 (Parseval); the forward
    # and inverse scales multiply to 1/N under every norm, so the inverse's bound is c log2(N) eps ||x||_2
    s_fwd = {0: 1.0, 1: 1.0 / points, 2: points**-0.5}[norm]
    c = (8 + 4 * np.log2(points + 1)) * eps * np.linalg.norm(xr)
    ref = np.fft.rfftn(xr, axes=axes, norm=npn)
    got = spec.cpu().numpy().astype(np.complex128)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= c * np.sqrt(points) * s_fwd
    assert torch.equal(spec, spec_in)  # c2r leaves its input unchanged
    sizes = [shape[a] for a in axes]
    ref_back = np.fft.irfftn(got, s=sizes, axes=axes, norm=npn)
    out = back.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(out - ref_back)) <= c
    assert np.max(np.abs(out - xr)) <= 2 * c
