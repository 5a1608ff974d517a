"""aether arrays on MI355X: the reference's Python package ``aether`` (dsc/aether/__init__.py:1-161) over
this library's ``aether-memory/1`` seam.

Use it as ``import aether_amd.dsc as aether``. One ``Array`` type for every place memory can live:
``device="cpu"``, ``"pinned"`` (page-locked host memory), ``"cuda"`` or ``"cuda:N"`` (HIP devices, named
as torch names them). Host and pinned arrays speak the buffer protocol (``memoryview``, ``numpy.asarray``
see them without a copy); every array speaks DLPack; device arrays also expose
``__cuda_array_interface__``. Importing needs neither torch, numpy nor a GPU: the device library
(``libaether_hip.so``) is opened on the first device call, and without a usable GPU that call raises
``CudaUnavailable``. ``aether_amd.dsc.fft`` is the reference's ``aether.fft`` over the same arrays.
"""
import atexit as _atexit

from . import _core
from ._core import Array, CudaUnavailable, Error, FftUnavailable, cuda_plugin, cuda_runtime_version
from ._dtype import Device, DType
from ._dtype import parse as _dtype

_core._register(Device, DType)
_atexit.register(_core._drain)  # frees a capture deferred, at the latest at exit

__all__ = ["fft", "Array", "Device", "DType", "Error", "CudaUnavailable", "FftUnavailable", "asarray", "copy",
           "cuda_plugin", "cuda_runtime_version", "empty", "empty_like", "from_dlpack", "full", "pinned", "zeros",
           "zeros_like"]

_KINDS = {"cpu": 0, "cuda": 1, "pinned": 2}


def _device(device):
    """(kind, index) of "cpu", "pinned", "cuda" (the current device), "cuda:N" or a Device."""
    if isinstance(device, Device):
        return _KINDS[device.kind], device.index
    if device == "cpu" or device == "pinned":
        return _KINDS[device], 0
    if device == "cuda":
        return 1, _core._current_device()
    if isinstance(device, str) and device.startswith("cuda:") and device[5:].isdigit():
        return 1, int(device[5:])
    raise Error(f"device= wants 'cpu', 'pinned', 'cuda' or 'cuda:N', not {device!r}")


def _shape(shape):
    if isinstance(shape, int):
        return [shape]
    return [int(s) for s in shape]


def _new(shape, dtype, device, stream, zero):
    kind, index = _device(device)
    return _core._new(kind, index, _shape(shape), _dtype(dtype).name, stream, zero)


def empty(shape, dtype="float64", device="cpu", stream=None):
    """An uninitialised array (device memory is stream-ordered on ``stream``)."""
    return _new(shape, dtype, device, stream, False)


def zeros(shape, dtype="float64", device="cpu", stream=None):
    return _new(shape, dtype, device, stream, True)


def full(shape, value, dtype="float64", device="cpu", stream=None):
    return _new(shape, dtype, device, stream, False).fill(value)


def pinned(shape, dtype="float64"):
    """Page-locked host memory, uninitialised: copies to and from a device are direct DMA."""
    return empty(shape, dtype, "pinned")


def _like(x, dtype, device):
    if device is None:
        device = x.device if isinstance(x, Array) else "cpu"
    return x.shape, x.dtype if dtype is None else dtype, device


def empty_like(x, dtype=None, device=None, stream=None):
    shape, dt, dev = _like(x, dtype, device)
    return empty(shape, dt, dev, stream)


def zeros_like(x, dtype=None, device=None, stream=None):
    shape, dt, dev = _like(x, dtype, device)
    return zeros(shape, dt, dev, stream)


def from_dlpack(x, stream=None):
    """View a foreign array without a copy; the producer's deleter runs when the last view dies."""
    return _core._from_dlpack(x, stream)


def _view(x, stream):
    """x as an Array without a copy, or None when it is nothing aether can view."""
    if isinstance(x, Array):
        return x
    if hasattr(x, "__cuda_array_interface__"):  # asks no stream ordering: legal inside a capture
        return _core._from_cai(x)
    if hasattr(x, "__dlpack_device__"):
        return from_dlpack(x, stream)
    try:
        memoryview(x)
    except TypeError:
        return None
    return _core._from_buffer(x)


def _from_nested(x, dtype):
    values, shape = [], []

    def walk(v, depth):
        if isinstance(v, (list, tuple)):
            if len(shape) == depth:
                shape.append(len(v))
            elif shape[depth] != len(v):
                raise Error("asarray: a ragged nested list has no array shape")
            for e in v:
                walk(e, depth + 1)
        else:
            if depth != len(shape):
                raise Error("asarray: a ragged nested list has no array shape")
            values.append(v)

    walk(x, 0)
    if dtype is None:
        if any(isinstance(v, float) for v in values):
            dtype = "float64"
        elif values and all(isinstance(v, bool) for v in values):
            dtype = "bool"
        else:
            dtype = "int64"
    out = empty(shape, dtype, "cpu")
    for i, v in enumerate(values):
        out._set_flat(i, v)
    return out


def asarray(x, dtype=None, device=None, stream=None):
    """x on ``device``: without a copy when it is already there, copied otherwise (``device=None`` keeps
    it where it is). Takes an Array, anything with the buffer protocol, DLPack or
    ``__cuda_array_interface__``, or a (nested) list of numbers. aether does no casting."""
    src = _view(x, stream)
    if src is None:
        src = _from_nested(x, dtype)
    if dtype is not None and src.dtype != _dtype(dtype):
        raise Error(f"asarray: the data is {src.dtype}; aether does no casting (asked for {_dtype(dtype)})")
    if device is None:
        return src
    kind, index = _device(device)
    here = src.device
    if _KINDS[here.kind] == kind and (kind != 1 or here.index == index):
        return src
    out = _core._new(kind, index, list(src.shape), src.dtype.name, stream, False)
    return copy(src, into=out, stream=stream)


def copy(src, into=None, stream=None):
    """Copy src into ``into`` (a new array on src's device when omitted); the direction is inferred.
    ``stream=None``: synchronous; otherwise queued on ``stream`` (a copy node inside a capture)."""
    s = _view(src, stream)
    if s is None:
        raise Error("copy: src is not an array aether can view")
    if into is None:
        into = empty_like(s, stream=stream)
    d = _view(into, None)
    if d is None:
        raise Error("copy: into is not an array aether can view")
    _core._copy(s, d, stream)
    return into


from . import fft  # noqa: E402  (batched FFTs over these arrays)
