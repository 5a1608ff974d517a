"""aether arrays (aether_amd.dsc), host rows: the reference's acceptance rows for its Python array package
(dsc/tests/test_arrays.py) that need no GPU -- import without numpy / torch / a device, dtype spellings,
the added 16-bit and 8-bit dtypes, the buffer protocol, DLPack exports and imports with their deleters,
copies, asarray, threads. The GPU rows are in test_gpu_dsc_arrays.py.

16-bit conversions are checked against numpy's float32 -> float16 and torch's float32 -> bfloat16 casts
(both round to nearest even) over every class of input: normals, subnormal results, overflow, ties, NaN."""
import ctypes
import gc
import pathlib
import struct
import subprocess
import sys
import textwrap
import threading

import numpy as np
import pytest

import aether_amd.dsc as aether
from aether_amd.dsc import _core

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "aether_amd" / "dsc"


def _run(code, **env):
    import os

    full = dict(os.environ, **env)
    full["PYTHONPATH"] = str(ROOT)
    return subprocess.run([sys.executable, "-W", "ignore", "-c", textwrap.dedent(code)], capture_output=True,
                          text=True, timeout=180, env=full)


def test_import_without_numpy_torch_or_a_device():
    out = _run("""
        import sys
        for m in ('numpy', 'cupy', 'torch'):
            sys.modules[m] = None
        import aether_amd.dsc as aether
        a = aether.zeros(4)
        assert aether.asarray([1.0, 2.0]).shape == (2,)
        assert 'torch' not in sys.modules or sys.modules['torch'] is None
        try:
            aether.empty(8, device='cuda')
        except aether.CudaUnavailable:
            print('ok')
    """, HIP_VISIBLE_DEVICES="-1")
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr[-600:]


def test_no_numpy_or_torch_import_in_the_package_sources():
    for path in list(PKG.glob("*.py")) + [PKG / "_core.cpp"]:
        text = path.read_text()
        for lib in ("numpy", "torch"):
            assert f"import {lib}" not in text and f"from {lib}" not in text, path


def test_dtype_spellings_and_refusals():
    assert aether.zeros(2, "float32").dtype == "float32"
    assert aether.zeros(2, float).dtype == "float64" and aether.zeros(2, int).dtype.name == "int64"
    assert aether.zeros(2, bool).dtype == "bool"
    assert aether.zeros(2, np.float32).dtype == "float32" and aether.zeros(2, np.dtype("int16")).dtype == "int16"
    assert aether.zeros(2, np.bool_).dtype == "bool"
    for bad in ("float128", "int4", "complex64", "complex128", "str", object):
        with pytest.raises(aether.Error):
            aether.empty(4, bad)
    assert aether.DType("float16").itemsize == 2 and aether.DType("bfloat16").code == 4


_NEW_DTYPES = [
    ("int8", -5, b"\xfb"), ("int16", -300, (-300).to_bytes(2, "little", signed=True)),
    ("uint16", 60000, (60000).to_bytes(2, "little")), ("float16", 1.5, struct.pack("<e", 1.5)),
    ("bfloat16", 1.0, b"\x80\x3f"),
]


def _bytes_of(a):
    return ctypes.string_at(a.ptr, a.nbytes)


@pytest.mark.parametrize("name,value,raw", _NEW_DTYPES, ids=[d[0] for d in _NEW_DTYPES])
def test_added_dtypes_allocate_fill_and_roundtrip_through_dlpack(name, value, raw):
    a = aether.full((3, 2), value, name)
    assert a.dtype == name and a.nbytes == 6 * len(raw) and a.dtype.itemsize == len(raw)
    assert _bytes_of(a) == raw * 6
    assert _bytes_of(aether.zeros(4, name)) == bytes(4 * len(raw))
    b = aether.from_dlpack(a)  # versioned capsule, and the legacy one below
    assert (b.ptr, b.shape, b.dtype) == (a.ptr, a.shape, a.dtype) and _bytes_of(b) == raw * 6
    legacy = aether.from_dlpack(type("L", (), {
        "__dlpack_device__": a.__dlpack_device__,
        "__dlpack__": lambda self, **kw: a.__dlpack__(**{k: v for k, v in kw.items() if k != "max_version"})})())
    assert legacy.ptr == a.ptr and legacy.dtype == name


def test_added_dtype_buffer_protocol():
    half = memoryview(aether.full(3, 1.5, "float16"))
    assert (half.format, half.itemsize, half.nbytes) == ("e", 2, 6)
    assert np.asarray(aether.full(3, 1.5, "float16")).tolist() == [1.5] * 3
    assert memoryview(aether.full(3, -4, "int8")).tolist() == [-4] * 3
    with pytest.raises(BufferError):
        memoryview(aether.zeros(2, "bfloat16"))


def test_half_fill_matches_numpy_rounding():
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.standard_normal(300).astype(np.float32) * np.float32(1000),
        (rng.uniform(-1, 1, 300) * 2.0 ** rng.integers(-30, -13, 300)).astype(np.float32),  # half subnormals
        np.array([65504.0, 65519.0, 65520.0, 1e6, -1e6, 2.0 ** -25, 2.0 ** -24, 1.5 * 2.0 ** -25, 2.0 ** -26,
                  1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 0.0, -0.0, np.inf, -np.inf, 1e-40], np.float32),
    ])
    for v in vals:
        got = struct.unpack("<H", _bytes_of(aether.full(1, float(v), "float16")))[0]
        with np.errstate(over="ignore"):
            want = int(np.array([v], np.float32).astype(np.float16).view(np.uint16)[0])
        assert got == want, (float(v), hex(got), hex(want))
    nan = struct.unpack("<H", _bytes_of(aether.full(1, float("nan"), "float16")))[0]
    assert (nan & 0x7C00) == 0x7C00 and (nan & 0x3FF)


def test_bfloat16_fill_matches_torch_rounding():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    vals = np.concatenate([rng.standard_normal(500).astype(np.float32) * np.float32(1e3),
                           np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.4e38, -3.4e38, 1e-40, 0.0, np.inf],
                                    np.float32)])
    want = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    for v, w in zip(vals, want):
        got = struct.unpack("<H", _bytes_of(aether.full(1, float(v), "bfloat16")))[0]
        assert got == int(w), (float(v), hex(got), hex(int(w)))


def test_host_surface_and_values():
    a = aether.zeros((3, 5), "float32")
    assert (a.shape, a.strides, a.nbytes, a.ndim, a.is_contiguous) == ((3, 5), (20, 4), 60, 2, True)
    assert a.device == aether.Device("cpu", 0) and a.dtype.itemsize == 4 and a.size == 15
    for dt, val in (("float64", 0.1), ("float32", 2.5), ("int64", -7), ("int32", 7), ("uint64", 9), ("uint32", 8),
                    ("uint8", 200), ("bool", True)):
        f = aether.full(33, val, dt)
        assert all(x == val or abs(x - val) < 1e-6 for x in memoryview(f).tolist()), dt
    assert aether.empty(0).ptr == 0 and aether.empty((0, 3)).nbytes == 0
    assert aether.asarray([[1, 2], [3, 4]]).dtype == "int64"
    assert aether.asarray([True, False]).dtype == "bool"
    assert repr(a) == "aether.Array(shape=(3, 5), dtype=float32, device='cpu')"
    with pytest.raises(aether.Error, match="rank"):
        aether.zeros((1,) * 7)
    with pytest.raises(aether.Error, match="ragged"):
        aether.asarray([[1.0, 2.0], [3.0]])


def test_copy_and_asarray_on_the_host():
    src = aether.asarray([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    dst = aether.copy(src)
    assert memoryview(dst).tolist() == memoryview(src).tolist() and dst.ptr != src.ptr
    assert aether.asarray(src) is src
    with pytest.raises(aether.Error):
        aether.copy(src, into=aether.zeros(5))
    with pytest.raises(aether.Error, match="no casting"):
        aether.asarray(src, dtype="float32")
    strided = aether.asarray(np.arange(12.0).reshape(3, 4)[:, ::2])
    assert strided.shape == (3, 2) and strided.strides == (32, 16) and not strided.is_contiguous
    with pytest.raises(aether.Error, match="contiguous"):
        aether.copy(strided)
    assert np.asarray(strided).tolist() == [[0.0, 2.0], [4.0, 6.0], [8.0, 10.0]]


def test_numpy_torch_memoryview_zero_copy():
    a = aether.zeros((3, 7), "float64")
    m = memoryview(a)
    m[1, 2] = 4.0
    assert memoryview(a)[1, 2] == 4.0
    v = np.asarray(a)
    assert v.ctypes.data == a.ptr and v[1, 2] == 4.0
    v[0, 0] = 9.0
    assert memoryview(a)[0, 0] == 9.0
    assert np.from_dlpack(a).ctypes.data == a.ptr
    torch = pytest.importorskip("torch")
    t = torch.from_dlpack(a)
    assert t.data_ptr() == a.ptr and float(t[0, 0]) == 9.0
    back = aether.asarray(np.arange(6.0).reshape(2, 3))
    assert back.shape == (2, 3) and memoryview(back).tolist()[1] == [3.0, 4.0, 5.0]
    tt = aether.asarray(torch.arange(4, dtype=torch.int32))
    assert tt.dtype == "int32" and memoryview(tt).tolist() == [0, 1, 2, 3]


def test_exports_release_exactly_once_on_the_host():
    before = _core._exports_live()
    a = aether.zeros(16, "float32")
    cap_v, cap_l = a.__dlpack__(max_version=(1, 0)), a.__dlpack__()
    assert _core._exports_live() == before + 2
    del cap_v, cap_l  # never consumed: the capsule destructor releases
    gc.collect()
    assert _core._exports_live() == before
    view = np.from_dlpack(a)  # consumed: the consumer's deleter releases
    assert _core._exports_live() == before + 1
    del a  # the export keeps the memory alive
    gc.collect()
    view[3] = 2.0
    assert float(view.sum()) == 2.0
    del view
    gc.collect()
    assert _core._exports_live() == before


def test_import_deleter_fires_once_and_keeps_the_producer_alive():
    x = np.arange(8.0)
    a = aether.from_dlpack(x)
    assert a.ptr == x.ctypes.data
    del x
    gc.collect()
    assert memoryview(a).tolist() == [float(i) for i in range(8)]
    del a
    gc.collect()


def test_dlpack_arguments():
    a = aether.zeros(4)
    with pytest.raises(BufferError):
        a.__dlpack__(copy=True)
    with pytest.raises(BufferError):
        a.__dlpack__(dl_device=(10, 0))
    assert a.__dlpack_device__() == (1, 0)
    a.__dlpack__(dl_device=(1, 0), copy=False, stream=None)


def test_threads_allocate_free_export():
    errors, before = [], _core._exports_live()
    gate = threading.Barrier(8)

    def work():
        try:
            gate.wait()
            for _ in range(50):
                a = aether.zeros(256, "float32")
                cap = a.__dlpack__(max_version=(1, 0))
                b = aether.full(16, 3, "int32")
                del cap, a, b
        except BaseException as exc:  # noqa: BLE001
            errors.append(exc)

    pool = [threading.Thread(target=work) for _ in range(8)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    gc.collect()
    assert not errors and _core._exports_live() == before


def test_core_size():
    assert pathlib.Path(_core.__file__).stat().st_size <= 1_500_000


def test_the_retired_plugin_major_override_is_a_named_error():
    out = _run("""
        import aether_amd.dsc as aether
        try:
            aether.empty(4, device='cuda')
        except aether.Error as exc:
            print('AETHER_CUDA_MAJOR is no longer used' in str(exc))
    """, AETHER_CUDA_MAJOR="13")
    assert out.returncode == 0 and out.stdout.strip() == "True", out.stderr[-300:]


def test_fft_misuse_is_a_named_error_before_any_device_call():
    F = aether.fft
    x = aether.zeros((8, 2), "float32")
    with pytest.raises(aether.Error, match="trailing axis of 2"):
        F.fft(aether.zeros((8,), "float32"))
    with pytest.raises(aether.Error, match="norm"):
        F.fft(x, norm="bogus")
    with pytest.raises(aether.Error, match="out of range"):
        F.fft(x, axis=3)
    with pytest.raises(aether.Error, match="kind"):
        F.plan("c3c", "float32", (4,), (0,))
    with pytest.raises(aether.Error, match="irfft"):
        F.irfft(aether.zeros((5, 2)), n=20)
    with pytest.raises(aether.Error, match="float32 or float64"):
        F.plan("c2c", "int32", (4,), (0,))
