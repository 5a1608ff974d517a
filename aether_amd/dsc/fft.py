"""Batched FFTs on aether arrays: the reference's ``aether.fft`` (dsc/aether/fft.py:1-148) on MI355X.

Complex data is a real float32 / float64 array whose LAST axis has extent 2 (real, imaginary); every
``axis`` counts the data axes only. Normalisation is numpy's ``norm``: "backward" (forward unscaled, inverse
scaled by 1/n), "forward" or "ortho". Every transform runs on the GPU through the seam's plans (Stockham
up to 8192 points, four-step up to 2^26, Bluestein for other lengths up to 2^25): device arrays in place on
their device; host and pinned arrays are staged through device memory and the result is a host array.
Lengths beyond those raise ``FftUnavailable``.

Plans are built eagerly, outside a stream capture; ``plan(...)`` (and the first transform of a new shape)
builds one, later calls reuse it. A device plan's ``execute(input, output, stream=None)`` allocates nothing
and may run inside a capture; ``stream=None`` runs on the input's allocation stream.
"""
from . import _core
from ._core import Error

__all__ = ["fft", "ifft", "rfft", "irfft", "fftn", "ifftn", "rfftn", "irfftn", "plan"]

_KIND = {"c2c": 0, "r2c": 1, "c2r": 2}
_NORM = {"backward": 0, "forward": 1, "ortho": 2}
_plans = {}


def _norm_code(norm):
    try:
        return _NORM[norm]
    except (KeyError, TypeError):
        raise Error(f"norm= wants 'backward', 'forward' or 'ortho', not {norm!r}") from None


def plan(kind, dtype, shape, axes, forward=True, norm="backward", device="cpu", stream=None):
    """A transform plan. ``kind``: "c2c", "r2c" or "c2r". ``shape`` is the data shape without the
    trailing 2: the complex shape for c2c, the REAL side's for r2c / c2r (whose complex side has
    ``shape[axes[-1]] // 2 + 1`` points along the last of ``axes``). ``forward`` applies to c2c only.
    Cached per (kind, dtype, shape, axes, direction, norm, device)."""
    from . import _device, _dtype

    if kind not in _KIND:
        raise Error(f"kind= wants 'c2c', 'r2c' or 'c2r', not {kind!r}")
    shape = tuple(int(s) for s in shape)
    axes = tuple(int(a) % len(shape) for a in axes) if shape else tuple(int(a) for a in axes)
    dkind, dindex = _device(device)
    if dkind == 2:  # pinned memory is host memory
        dkind = 0
    name = _dtype(dtype).name
    key = (kind, name, shape, axes, bool(forward), norm, dkind, dindex)
    p = _plans.get(key)
    if p is None:
        p = _core._fft_plan(_KIND[kind], name, dkind, dindex, list(shape), list(axes), bool(forward),
                            _norm_code(norm), stream)
        _plans[key] = p
    return p


def _axes(axis, rank):
    if axis is None:
        axis = tuple(range(rank))
    elif isinstance(axis, int):
        axis = (axis,)
    out = []
    for a in axis:
        if not -rank <= a < rank:
            raise Error(f"axis {a} is out of range for {rank} data axes")
        out.append(a % rank)
    return tuple(out)


def _where(x):
    d = x.device
    return f"cuda:{d.index}" if d.kind == "cuda" else "cpu"


def _complex_shape(x):
    if x.ndim < 2 or x.shape[-1] != 2:
        raise Error("a complex array is a float array with a trailing axis of 2 (re, im); this one has shape "
                    f"{tuple(x.shape)}")
    return tuple(x.shape[:-1])


def _run(kind, x, out_shape, shape, axes, forward, norm, stream):
    from . import empty

    where = _where(x)
    p = plan(kind, x.dtype, shape, axes, forward, norm, where, stream)
    out = empty(out_shape, x.dtype, where, stream)
    p.execute(x, out, stream)
    return out


def _c2c(x, axes, forward, norm, stream):
    shape = _complex_shape(x)
    _norm_code(norm)
    return _run("c2c", x, shape + (2,), shape, _axes(axes, len(shape)), forward, norm, stream)


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
    """N-D transform of REAL data: the output keeps n // 2 + 1 points along the last of ``axes``."""
    shape = tuple(x.shape)
    ax = _axes(axes, len(shape))
    half = list(shape)
    half[ax[-1]] = shape[ax[-1]] // 2 + 1
    return _run("r2c", x, tuple(half) + (2,), shape, ax, True, norm, stream)


def rfft(x, axis=-1, norm="backward", stream=None):
    return rfftn(x, (axis,), norm, stream)


def irfftn(x, s=None, axes=None, norm="backward", stream=None):
    """Inverse of ``rfftn``. ``s``: the REAL output's lengths along ``axes``; by default the last axis
    gets the even length 2 (m - 1), as numpy does."""
    shape = _complex_shape(x)
    ax = _axes(axes, len(shape))
    real = list(shape)
    if s is None:
        real[ax[-1]] = 2 * (shape[ax[-1]] - 1)
    else:
        for a, n in zip(ax, s):
            real[a] = int(n)
    if real[ax[-1]] // 2 + 1 != shape[ax[-1]]:
        raise Error(f"irfft: a real length of {real[ax[-1]]} has {real[ax[-1]] // 2 + 1} points, "
                    f"not {shape[ax[-1]]}")
    return _run("c2r", x, tuple(real), tuple(real), ax, False, norm, stream)


def irfft(x, n=None, axis=-1, norm="backward", stream=None):
    return irfftn(x, None if n is None else (n,), (axis,), norm, stream)
