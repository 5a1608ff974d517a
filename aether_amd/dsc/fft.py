"""Transforms of aether arrays on MI355X: the public surface of the reference's ``aether.fft``
(dsc/aether/fft.py) -- ``fft``, ``ifft``, ``fftn``, ``ifftn``, ``rfft``, ``irfft``, ``rfftn``, ``irfftn`` and
``plan`` -- over the seam's FFT plans.

Conventions (the interface): complex data are float32 / float64 arrays whose last axis holds (re, im), and
``axis`` / ``axes`` count the other ("data") axes; ``norm`` is numpy's ("backward", "forward", "ortho").
Every transform runs on the GPU: device arrays on their device, host and pinned arrays staged through it
(the result is then a host array). Power-of-two lengths up to 2^34 and other lengths below 2^32 are
supported (device memory permitting); longer ones raise ``FftUnavailable``.

A plan is built once per (kind, dtype, shape, axes, direction, norm, place) and kept; building one is
refused inside a stream capture, while a device plan's ``execute(input, output, stream=None)`` allocates
nothing and may be captured (``stream=None``: the input's allocation stream).
"""
from . import _core
from ._core import Error
from ._shapes import KINDS, NORMS, complex_data, data_axes, real_lengths, seam_code, spectrum

__all__ = ["fft", "ifft", "rfft", "irfft", "fftn", "ifftn", "rfftn", "irfftn", "plan"]

_cache = {}


def plan(kind, dtype, shape, axes, forward=True, norm="backward", device="cpu", stream=None):
    """The plan of one transform. ``kind`` is "c2c", "r2c" or "c2r"; ``shape`` the data shape (no
    trailing 2): the complex one for c2c, the REAL side's for r2c / c2r, whose complex side has
    ``shape[axes[-1]] // 2 + 1`` points along the last of ``axes``. ``forward`` matters for c2c only.
    Equal requests return the same plan object."""
    from . import _place, _resolve_dtype

    kind_code = seam_code(KINDS, kind, "kind", Error)
    norm_code = seam_code(NORMS, norm, "norm", Error)
    dims = tuple(int(e) for e in shape)
    ax = tuple(int(a) % len(dims) for a in axes) if dims else tuple(int(a) for a in axes)
    where = _place(device)
    place = (1, where.index) if where.kind == "cuda" else (0, 0)  # pinned memory is host memory
    name = _resolve_dtype(dtype).name
    key = (kind_code, name, dims, ax, bool(forward), norm_code) + place
    found = _cache.get(key)
    if found is None:
        found = _core._fft_plan(kind_code, name, place[0], place[1], list(dims), list(ax), bool(forward),
                                norm_code, stream)
        _cache[key] = found
    return found


def _execute(kind, x, shape, axes, forward, norm, out_shape, stream):
    from . import empty

    where = "cpu" if x.device.kind != "cuda" else f"cuda:{x.device.index}"
    p = plan(kind, x.dtype, shape, axes, forward, norm, where, stream)
    out = empty(out_shape, x.dtype, where, stream)
    p.execute(x, out, stream)
    return out


def _complex(x, axes, forward, norm, stream):
    seam_code(NORMS, norm, "norm", Error)
    dims = complex_data(x.shape, Error)
    return _execute("c2c", x, dims, data_axes(axes, len(dims), Error), forward, norm, tuple(x.shape), stream)


def fft(x, axis=-1, norm="backward", stream=None):
    """Transform of a complex array along one data axis."""
    return _complex(x, (axis,), True, norm, stream)


def ifft(x, axis=-1, norm="backward", stream=None):
    """Inverse transform along one data axis."""
    return _complex(x, (axis,), False, norm, stream)


def fftn(x, axes=None, norm="backward", stream=None):
    """Transform over several data axes (default: all)."""
    return _complex(x, axes, True, norm, stream)


def ifftn(x, axes=None, norm="backward", stream=None):
    """Inverse transform over several data axes (default: all)."""
    return _complex(x, axes, False, norm, stream)


def rfftn(x, axes=None, norm="backward", stream=None):
    """Transform of REAL data over ``axes``; n // 2 + 1 points are kept along the last of them."""
    dims = tuple(x.shape)
    ax = data_axes(axes, len(dims), Error)
    return _execute("r2c", x, dims, ax, True, norm, spectrum(dims, ax) + (2,), stream)


def rfft(x, axis=-1, norm="backward", stream=None):
    """Transform of REAL data along one axis."""
    return rfftn(x, (axis,), norm, stream)


def irfftn(x, s=None, axes=None, norm="backward", stream=None):
    """Inverse of ``rfftn``: REAL output. ``s`` gives the real lengths along ``axes``; without it the last
    of them becomes 2 (m - 1) for m complex points (numpy's default)."""
    dims = complex_data(x.shape, Error)
    ax = data_axes(axes, len(dims), Error)
    real = real_lengths(dims, ax, s, Error)
    return _execute("c2r", x, real, ax, False, norm, real, stream)


def irfft(x, n=None, axis=-1, norm="backward", stream=None):
    """Inverse of ``rfft`` along one axis; ``n`` is the real output length."""
    return irfftn(x, None if n is None else (n,), (axis,), norm, stream)
