"""aether arrays (aether_amd.dsc) on the host -- no GPU.

The acceptance rows are those of the reference package's host tests (dsc/tests/test_arrays.py), restated:
the package imports without numpy, torch or a device and refuses device memory by name; dtypes are
spelled every way the interface allows and refused otherwise; the 8- and 16-bit types allocate, fill, cross
DLPack and the buffer protocol; host arrays share memory with memoryview / numpy / torch; DLPack exports
and imports release exactly once, also from many threads; copy/asarray move, never cast; FFT misuse is a
named error before any device call. 16-bit fills are checked against numpy's float32 -> float16 and
torch's float32 -> bfloat16 casts (round to nearest even) over normals, subnormal results, overflow,
ties and NaN."""
import ctypes
import gc
import os
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

REPO = pathlib.Path(__file__).resolve().parents[1]


def python_child(source, **env):
    """Run `source` in a fresh interpreter with the repository importable; returns CompletedProcess."""
    environ = {**os.environ, **env, "PYTHONPATH": str(REPO)}
    return subprocess.run([sys.executable, "-W", "ignore", "-c", textwrap.dedent(source)], env=environ,
                          capture_output=True, text=True, timeout=180)


def raw(a):
    """The bytes of a host array."""
    return ctypes.string_at(a.ptr, a.nbytes)


def as_half_bits(x):
    return struct.unpack("<H", raw(aether.full(1, float(x), "float16")))[0]


def as_bf16_bits(x):
    return struct.unpack("<H", raw(aether.full(1, float(x), "bfloat16")))[0]


# ---------------------------------------------------------------------------------------------- import
class TestImport:
    def test_without_numpy_torch_or_a_gpu(self):
        child = python_child("""
            import sys
            sys.modules.update(numpy=None, cupy=None, torch=None)
            import aether_amd.dsc as aether
            assert aether.zeros(3).shape == (3,) and aether.asarray([2.0, 3.0]).shape == (2,)
            try:
                aether.empty(4, device="cuda")
                print("allocated")
            except aether.CudaUnavailable:
                print("refused")
        """, HIP_VISIBLE_DEVICES="-1")
        assert child.returncode == 0 and child.stdout.split() == ["refused"], child.stderr[-800:]

    def test_package_sources_import_neither_numpy_nor_torch(self):
        sources = sorted((REPO / "aether_amd" / "dsc").glob("*.py")) + [REPO / "aether_amd" / "dsc" / "_core.cpp"]
        for path in sources:
            text = path.read_text()
            for mod in ("numpy", "torch"):
                assert f"import {mod}" not in text and f"from {mod}" not in text, (path.name, mod)

    def test_retired_plugin_major_override_names_itself(self):
        child = python_child("""
            import aether_amd.dsc as aether
            try:
                aether.empty(4, device="cuda")
            except aether.Error as exc:
                print("AETHER_CUDA_MAJOR is no longer used" in str(exc))
        """, AETHER_CUDA_MAJOR="12")
        assert child.returncode == 0 and child.stdout.strip() == "True", child.stderr[-400:]

    def test_extension_stays_small(self):
        assert pathlib.Path(_core.__file__).stat().st_size <= 1_500_000


# ---------------------------------------------------------------------------------------------- dtypes
SPELLINGS = [("float32", "float32"), (float, "float64"), (int, "int64"), (bool, "bool"), (np.float32, "float32"),
             (np.dtype("int16"), "int16"), (np.bool_, "bool"), (np.dtype(np.uint8), "uint8"),
             (aether.DType("bfloat16"), "bfloat16")]


class TestDtypes:
    @pytest.mark.parametrize("spelling,name", SPELLINGS, ids=[n for _, n in SPELLINGS])
    def test_spellings(self, spelling, name):
        a = aether.zeros(2, spelling)
        assert a.dtype == name and a.dtype.name == name and a.dtype is aether.DType(name)

    def test_torch_spelling_without_importing_it_here(self):
        torch = pytest.importorskip("torch")
        assert aether.zeros(2, torch.float16).dtype == "float16"

    @pytest.mark.parametrize("bad", ["float128", "int4", "complex64", "complex128", "str", object])
    def test_refusals(self, bad):
        with pytest.raises(aether.Error, match="not supported"):
            aether.empty(4, bad)

    def test_codes_and_widths(self):
        table = {"float16": (2, 2), "bfloat16": (4, 2), "int8": (0, 1), "uint64": (1, 8), "bool": (6, 1)}
        for name, (code, size) in table.items():
            t = aether.DType(name)
            assert (t.code, t.itemsize, t.bits) == (code, size, 8 * size)


# the 8- and 16-bit types: (name, fill value, expected bytes of one element)
SMALL = [("int8", -5, struct.pack("<b", -5)), ("int16", -300, struct.pack("<h", -300)),
         ("uint16", 60000, struct.pack("<H", 60000)), ("float16", 1.5, struct.pack("<e", 1.5)),
         ("bfloat16", 1.0, struct.pack("<H", 0x3F80))]


class TestSmallTypes:
    @pytest.mark.parametrize("name,value,elem", SMALL, ids=[s[0] for s in SMALL])
    def test_fill_zero_and_dlpack_both_generations(self, name, value, elem):
        a = aether.full((2, 3), value, name)
        assert a.dtype.itemsize == len(elem) and a.nbytes == 6 * len(elem) and raw(a) == elem * 6
        assert raw(aether.zeros(5, name)) == bytes(5 * len(elem))
        versioned = aether.from_dlpack(a)
        assert (versioned.ptr, versioned.shape, versioned.dtype) == (a.ptr, a.shape, a.dtype)

        class LegacyProducer:  # offers only the unversioned capsule
            def __dlpack_device__(self):
                return a.__dlpack_device__()

            def __dlpack__(self, stream=None, **ignored):
                if "max_version" in ignored:
                    raise TypeError("no max_version here")
                return a.__dlpack__(stream=stream)

        legacy = aether.from_dlpack(LegacyProducer())
        assert legacy.ptr == a.ptr and legacy.dtype == name and raw(legacy) == elem * 6

    def test_buffer_protocol(self):
        h = memoryview(aether.full(4, 1.5, "float16"))
        assert (h.format, h.itemsize, h.nbytes) == ("e", 2, 8)
        assert np.asarray(aether.full(4, 1.5, "float16")).tolist() == [1.5] * 4
        assert memoryview(aether.full(3, -4, "int8")).tolist() == [-4, -4, -4]
        with pytest.raises(BufferError):  # no struct format for bfloat16
            memoryview(aether.zeros(2, "bfloat16"))

    def test_float16_fill_rounds_like_numpy(self):
        rng = np.random.default_rng(15)
        cases = np.concatenate([
            (rng.standard_normal(250) * 800).astype(np.float32),
            (rng.uniform(-1, 1, 250) * np.exp2(rng.integers(-30, -13, 250))).astype(np.float32),  # subnormal halves
            np.float32([65504, 65519, 65520, 7e5, -7e5, 2**-24, 2**-25, 1.5 * 2**-25, 2**-26, 1 + 2**-11,
                        1 + 3 * 2**-11, 0.0, -0.0, np.inf, -np.inf, 1e-40]),
        ])
        with np.errstate(over="ignore"):
            want = cases.astype(np.float16).view(np.uint16)
        got = np.array([as_half_bits(v) for v in cases], np.uint16)
        assert np.array_equal(got, want), [(float(c), hex(g), hex(w)) for c, g, w in zip(cases, got, want) if g != w][:5]
        nan = as_half_bits(float("nan"))
        assert nan & 0x7C00 == 0x7C00 and nan & 0x03FF

    def test_bfloat16_fill_rounds_like_torch(self):
        torch = pytest.importorskip("torch")
        rng = np.random.default_rng(16)
        cases = np.concatenate([(rng.standard_normal(400) * 1e3).astype(np.float32),
                                np.float32([1 + 2**-8, 1 + 3 * 2**-8, 3.4e38, -3.4e38, 1e-40, 0.0, np.inf])])
        want = torch.from_numpy(cases).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
        got = np.array([as_bf16_bits(v) for v in cases], np.uint16)
        assert np.array_equal(got, want)


# ---------------------------------------------------------------------------------------------- host arrays
class TestHostArrays:
    def test_surface(self):
        a = aether.zeros((3, 5), "float32")
        assert (a.shape, a.strides, a.nbytes, a.ndim, a.size, a.is_contiguous) == ((3, 5), (20, 4), 60, 2, 15, True)
        assert a.device == aether.Device("cpu", 0) and str(a.device) == "cpu"
        assert repr(a) == "aether.Array(shape=(3, 5), dtype=float32, device='cpu')"
        assert aether.empty(0).ptr == 0 and aether.empty((0, 3)).nbytes == 0

    @pytest.mark.parametrize("dtype,value", [("float64", 0.1), ("float32", 2.5), ("int64", -7), ("int32", 7),
                                             ("uint64", 9), ("uint32", 8), ("uint8", 200), ("bool", True)])
    def test_full(self, dtype, value):
        got = np.asarray(aether.full(33, value, dtype))
        assert got.dtype == np.dtype(dtype) and np.all(got == np.array(value, dtype))

    def test_nested_lists(self):
        assert aether.asarray([[1, 2], [3, 4]]).dtype == "int64"
        assert aether.asarray([True, False]).dtype == "bool"
        assert aether.asarray([[1, 2.5]]).dtype == "float64"
        assert np.asarray(aether.asarray(((1.0, 2.0), (3.0, 4.0)))).tolist() == [[1.0, 2.0], [3.0, 4.0]]
        for ragged in ([[1.0, 2.0], [3.0]], [[1.0], 2.0], [1.0, [2.0]]):
            with pytest.raises(aether.Error, match="ragged"):
                aether.asarray(ragged)

    def test_rank_limit(self):
        with pytest.raises(aether.Error, match="rank"):
            aether.zeros((1,) * 7)

    def test_copy_and_asarray_move_never_cast(self):
        src = aether.asarray([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
        dup = aether.copy(src)
        assert dup.ptr != src.ptr and raw(dup) == raw(src)
        assert aether.asarray(src) is src and aether.asarray(src, device="cpu") is src
        with pytest.raises(aether.Error):
            aether.copy(src, into=aether.zeros(5))
        with pytest.raises(aether.Error, match="no casting"):
            aether.asarray(src, dtype="float32")
        assert aether.zeros_like(src).dtype == "float64" and aether.empty_like(src, dtype="int8").dtype == "int8"

    def test_strided_views(self):
        base = np.arange(12.0).reshape(3, 4)
        v = aether.asarray(base[:, ::2])
        assert v.shape == (3, 2) and v.strides == (32, 16) and not v.is_contiguous
        assert np.asarray(v).tolist() == [[0.0, 2.0], [4.0, 6.0], [8.0, 10.0]]
        with pytest.raises(aether.Error, match="contiguous"):
            aether.copy(v)

    def test_shared_memory_with_memoryview_numpy_torch(self):
        a = aether.zeros((3, 7))
        memoryview(a)[1, 2] = 4.0
        n = np.asarray(a)
        assert n.ctypes.data == a.ptr and n[1, 2] == 4.0
        n[0, 0] = 9.0
        assert memoryview(a)[0, 0] == 9.0 and np.from_dlpack(a).ctypes.data == a.ptr
        b = aether.asarray(np.arange(6.0).reshape(2, 3))
        assert memoryview(b).tolist() == [[0.0, 1.0, 2.0], [3.0, 4.0, 5.0]]
        torch = pytest.importorskip("torch")
        t = torch.from_dlpack(a)
        assert t.data_ptr() == a.ptr and t[0, 0].item() == 9.0
        ti = aether.asarray(torch.arange(4, dtype=torch.int32))
        assert ti.dtype == "int32" and memoryview(ti).tolist() == [0, 1, 2, 3]


# ---------------------------------------------------------------------------------------------- DLPack lifetimes
class TestDLPack:
    def test_each_export_is_released_once(self):
        live = _core._exports_live
        start = live()
        a = aether.zeros(16, "float32")
        caps = [a.__dlpack__(max_version=(1, 0)), a.__dlpack__()]
        assert live() == start + 2
        caps.clear()  # never consumed: released by the capsules' destructors
        gc.collect()
        assert live() == start
        consumer = np.from_dlpack(a)  # consumed: released by numpy's deleter call
        assert live() == start + 1
        del a  # the export keeps the memory alive
        gc.collect()
        consumer[5] = 3.0
        assert consumer.sum() == 3.0
        del consumer
        gc.collect()
        assert live() == start

    def test_imported_producer_stays_alive(self):
        x = np.linspace(0.0, 7.0, 8)
        a = aether.from_dlpack(x)
        assert a.ptr == x.ctypes.data
        del x
        gc.collect()
        assert memoryview(a).tolist() == [float(k) for k in range(8)]

    def test_export_arguments(self):
        a = aether.zeros(4)
        assert a.__dlpack_device__() == (1, 0)
        for bad in ({"copy": True}, {"dl_device": (10, 0)}):
            with pytest.raises(BufferError):
                a.__dlpack__(**bad)
        a.__dlpack__(dl_device=(1, 0), copy=False, stream=None)

    def test_threads(self):
        start, failures = _core._exports_live(), []
        barrier = threading.Barrier(8)

        def churn():
            try:
                barrier.wait()
                for _ in range(50):
                    x = aether.zeros(256, "float32")
                    cap = x.__dlpack__(max_version=(1, 0))
                    y = aether.full(16, 3, "int32")
                    del cap, x, y
            except BaseException as exc:  # noqa: BLE001
                failures.append(exc)

        threads = [threading.Thread(target=churn) for _ in range(8)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        gc.collect()
        assert not failures and _core._exports_live() == start


# ---------------------------------------------------------------------------------------------- FFT arguments
class TestFftArguments:
    """Misuse is an aether.Error raised before any device call (no GPU here)."""

    def test_named_errors(self):
        F = aether.fft
        pairs = aether.zeros((8, 2), "float32")
        checks = [
            (lambda: F.fft(aether.zeros(8, "float32")), "trailing axis of 2"),
            (lambda: F.fft(pairs, norm="unitary"), "norm"),
            (lambda: F.fft(pairs, axis=2), "out of range"),
            (lambda: F.fftn(aether.zeros((4, 4, 2)), axes=(0, -3)), "out of range"),
            (lambda: F.plan("r2r", "float32", (4,), (0,)), "kind"),
            (lambda: F.irfft(aether.zeros((5, 2)), n=20), "irfft"),
            (lambda: F.plan("c2c", "int32", (4,), (0,)), "float32 or float64"),
        ]
        for call, words in checks:
            with pytest.raises(aether.Error, match=words):
                call()
