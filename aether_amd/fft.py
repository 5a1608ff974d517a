"""Transforms of torch tensors with the reference's ``aether.fft`` API (dsc/aether/fft.py): the same
functions as ``aether_amd.dsc.fft``, for callers whose arrays are torch tensors.

Complex data are float32 / float64 tensors whose last axis holds (re, im); axes count the other axes;
``norm`` is numpy's ("backward", "forward", "ortho"); the shape rules are those of
``aether_amd.dsc._shapes``. Arrays are torch tensors (or anything torch can take by
DLPack, or numpy arrays); the transform always runs on the GPU through this library's seam plans
(``aether_memory_v1().fft_plan``): a host array is copied to the device and the result copied back.

Lengths: powers of two up to 8192 run the LDS-resident Stockham transform, powers of two up to 2^34 a
four-step transform over it (two levels past 2^26); other lengths below 2^32 run Bluestein's chirp-z
through the same transforms; anything longer raises ``FftUnavailable`` (code 2, as the reference's plugin does when its FFT
library is missing).

Plans are built eagerly, outside a stream capture; ``plan(...)`` (and the first call of a transform on a
new shape) builds one, later calls reuse it. ``Plan.execute(input, output, stream=None)`` allocates
nothing and may run inside a capture.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _seam
from ._lib import AetherError
from .dsc._shapes import NORMS, complex_data, data_axes, real_lengths, seam_code, spectrum

__all__ = ["fft", "ifft", "rfft", "irfft", "fftn", "ifftn", "rfftn", "irfftn", "plan", "Plan", "FftUnavailable",
           "Error"]

Error = AetherError
_KINDS = {"c2c": _seam.C2C, "r2c": _seam.R2C, "c2r": _seam.C2R}
_cache: dict = {}


def _error(message: str) -> AetherError:
    return AetherError(1, message)


class FftUnavailable(AetherError):
    """The transform is outside what this library's FFT plans cover (return code 2)."""

    def __init__(self, msg: str):
        super().__init__(2, msg)


def _norm(norm) -> int:
    return seam_code(NORMS, norm, "norm", _error)


def _dtype(dtype) -> torch.dtype:
    names = {"float32": torch.float32, "float64": torch.float64}
    if isinstance(dtype, torch.dtype) and dtype in (torch.float32, torch.float64):
        return dtype
    name = dtype if isinstance(dtype, str) else getattr(dtype, "__name__", None) or str(dtype)
    if name in names:
        return names[name]
    raise Error(1, f"FFTs take float32 or float64 data, not {dtype!r}")


def _device_index(device) -> int:
    if device is None or device == "cuda":
        return torch.cuda.current_device()
    if isinstance(device, torch.device):
        if device.type != "cuda":
            raise Error(1, f"plans run on a GPU; device {device} is not one")
        return device.index if device.index is not None else torch.cuda.current_device()
    if isinstance(device, str) and device.startswith("cuda:") and device[5:].isdigit():
        return int(device[5:])
    raise Error(1, f"device= wants 'cuda' or 'cuda:N', not {device!r}")


def _stream(stream) -> int:
    if stream is None:
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(getattr(stream, "cuda_stream", getattr(stream, "ptr", 0)))


class Plan:
    """A transform plan over the seam (dsc/src/cufft_plan.h:122-222): eager, reusable, capture-safe."""

    def __init__(self, kind: str, dtype: torch.dtype, shape: tuple, axes: tuple, forward: bool, norm: str,
                 device: int, stream):
        self.kind, self.dtype, self.shape, self.axes = kind, dtype, shape, axes
        self.forward = True if kind == "r2c" else False if kind == "c2r" else bool(forward)
        self.norm, self.device = norm, device
        half = spectrum(tuple(shape), axes) if kind != "c2c" and axes else tuple(shape)
        full, halfc = tuple(shape) + (2,), tuple(half) + (2,)
        self.input_shape = {"c2c": full, "r2c": tuple(shape), "c2r": halfc}[kind]
        self.output_shape = {"c2c": full, "r2c": halfc, "c2r": tuple(shape)}[kind]
        t = _seam.table()
        shp = (C.c_int64 * max(1, len(shape)))(*shape)
        ax = (C.c_int32 * max(1, len(axes)))(*axes)
        h = C.c_void_p()
        rc = t.fft_plan(device, _KINDS[kind], 1 if dtype == torch.float64 else 0, _norm(norm), int(self.forward),
                        len(shape), shp, len(axes), ax, C.c_void_p(_stream(stream)), C.byref(h))
        if rc == 2:
            raise FftUnavailable(_seam.last_error())
        if rc != 0:
            raise Error(rc, _seam.last_error())
        self._h = h

    def execute(self, input: torch.Tensor, output: torch.Tensor, stream=None) -> torch.Tensor:
        """Run on `stream` (default: torch's current stream): no allocation, legal inside a capture."""
        for name, x, want in (("input", input, self.input_shape), ("output", output, self.output_shape)):
            if x.dtype != self.dtype:
                raise Error(1, f"the plan wants {str(self.dtype).replace('torch.', '')}; the {name} is "
                               f"{str(x.dtype).replace('torch.', '')}")
            if tuple(x.shape) != tuple(want):
                raise Error(1, f"the {name} has the wrong shape {tuple(x.shape)} (the plan wants {tuple(want)})")
            if not x.is_cuda or x.device.index != self.device:
                raise Error(1, f"the {name} is not on cuda:{self.device}")
            if not x.is_contiguous():
                raise Error(1, f"the {name} is not dense row-major")
        if self.kind != "c2c" and input.data_ptr() == output.data_ptr() and input.numel():
            raise Error(1, "in == out is allowed for c2c only")
        rc = _seam.table().fft_exec(self._h, C.c_void_p(input.data_ptr()), C.c_void_p(output.data_ptr()),
                                    C.c_void_p(_stream(stream)))
        if rc != 0:
            raise Error(rc, _seam.last_error())
        return output


def plan(kind, dtype, shape, axes, forward=True, norm="backward", device="cuda", stream=None) -> Plan:
    """A plan (dsc/aether/fft.py:43-66). ``shape`` is the data shape without the trailing 2: the complex
    shape for c2c, the REAL side's for r2c / c2r. Cached per (kind, dtype, shape, axes, direction, norm,
    device)."""
    if kind not in _KINDS:
        raise _error(f"kind= must be one of 'c2c', 'r2c', 'c2r'; got {kind!r}")
    dt = _dtype(dtype)
    shape = tuple(int(s) for s in shape)
    axes = data_axes(axes, len(shape), _error)
    _norm(norm)
    dev = _device_index(device)
    fwd = True if kind == "r2c" else False if kind == "c2r" else bool(forward)
    key = (kind, dt, shape, axes, fwd, norm, dev)
    p = _cache.get(key)
    if p is None:
        p = Plan(kind, dt, shape, axes, fwd, norm, dev, stream)
        _cache[key] = p
    return p


def _as_tensor(x):
    """(device tensor, was_on_host) without a copy when `x` already lives on a GPU."""
    if not isinstance(x, torch.Tensor):
        if hasattr(x, "__dlpack__"):
            try:
                x = torch.from_dlpack(x)
            except Exception:
                x = torch.as_tensor(x)
        else:
            x = torch.as_tensor(x)
    host = not x.is_cuda
    if host:
        x = x.to("cuda")
    return x.contiguous(), host


def _run(kind, x, host, shape, axes, forward, norm, stream):
    p = plan(kind, x.dtype, shape, axes, forward, norm, x.device, stream)
    out = torch.empty(p.output_shape, dtype=x.dtype, device=x.device)
    p.execute(x, out, stream)
    return out.cpu() if host else out


def _c2c(x, axes, forward, norm, stream):
    x, host = _as_tensor(x)
    _dtype(x.dtype)
    shape = complex_data(x.shape, _error)
    return _run("c2c", x, host, shape, data_axes(axes, len(shape), _error), forward, norm, stream)


def fft(x, axis=-1, norm="backward", stream=None):
    """1-D transform of a complex array along ``axis``."""
    return _c2c(x, (axis,), True, norm, stream)


def ifft(x, axis=-1, norm="backward", stream=None):
    return _c2c(x, (axis,), False, norm, stream)


def fftn(x, axes=None, norm="backward", stream=None):
    """N-D transform over ``axes`` (default: every data axis)."""
    return _c2c(x, axes, True, norm, stream)


def ifftn(x, axes=None, norm="backward", stream=None):
    return _c2c(x, axes, False, norm, stream)


def rfftn(x, axes=None, norm="backward", stream=None):
    """N-D transform of REAL data: the output keeps ``n // 2 + 1`` points along the last of ``axes``."""
    x, host = _as_tensor(x)
    _dtype(x.dtype)
    shape = tuple(x.shape)
    return _run("r2c", x, host, shape, data_axes(axes, len(shape), _error), True, norm, stream)


def rfft(x, axis=-1, norm="backward", stream=None):
    return rfftn(x, (axis,), norm, stream)


def irfftn(x, s=None, axes=None, norm="backward", stream=None):
    """Inverse of ``rfftn``. ``s`` is the REAL output's shape along ``axes``; the default recovers an even
    length on the last axis (``2 * (m - 1)``), as numpy does. The input is left unchanged."""
    x, host = _as_tensor(x)
    _dtype(x.dtype)
    shape = complex_data(x.shape, _error)
    ax = data_axes(axes, len(shape), _error)
    return _run("c2r", x, host, real_lengths(shape, ax, s, _error), ax, False, norm, stream)


def irfft(x, n=None, axis=-1, norm="backward", stream=None):
    return irfftn(x, None if n is None else (n,), (axis,), norm, stream)
