"""aether arrays on MI355X: the public surface of the reference's Python package ``aether``
(dsc/aether/__init__.py) -- ``empty``, ``zeros``, ``full``, ``pinned``, ``*_like``, ``asarray``, ``copy``,
``from_dlpack``, ``Array``, ``Device``, ``DType`` and its errors -- over this library's ``aether-memory/1``
seam (``libaether_hip.so``).

Use it as ``import aether_amd.dsc as aether``. One ``Array`` type covers every place memory lives:
``device="cpu"``, ``"pinned"``, ``"cuda"`` (the current HIP device) or ``"cuda:N"``. Host and pinned
arrays speak the buffer protocol, every array speaks DLPack, device arrays also expose
``__cuda_array_interface__``. Importing needs neither torch, numpy nor a GPU: the device library is opened
by the first device call, which raises ``CudaUnavailable`` when no GPU can be used. ``aether_amd.dsc.fft``
holds the transforms.
"""
import atexit as _atexit

from . import _core
from ._core import Array, CudaUnavailable, Error, FftUnavailable, cuda_plugin, cuda_runtime_version
from ._types import Device, DType
from ._types import resolve as _resolve_dtype

_core._register(Device, DType)
# chunks released while a capture was running are freed by drain(); at the latest when Python exits
_atexit.register(_core._drain)

__all__ = ["fft", "Array", "Device", "DType", "Error", "CudaUnavailable", "FftUnavailable", "asarray", "copy",
           "cuda_plugin", "cuda_runtime_version", "empty", "empty_like", "from_dlpack", "full", "pinned", "zeros",
           "zeros_like"]


def _place(device):
    """The Device a ``device=`` argument names."""
    if isinstance(device, Device):
        return device
    if isinstance(device, str):
        kind, sep, ordinal = device.partition(":")
        if not sep and kind in ("cpu", "pinned"):
            return Device(kind, 0)
        if kind == "cuda" and not sep:
            return Device("cuda", _core._current_device())
        if kind == "cuda" and ordinal.isdigit():
            return Device("cuda", int(ordinal))
    raise Error(f"device= wants 'cpu', 'pinned', 'cuda' or 'cuda:N', not {device!r}")


def _extents(shape):
    return [shape] if isinstance(shape, int) else [int(e) for e in shape]


def _allocate(shape, dtype, device, stream, zero):
    where = _place(device)
    return _core._new(where.seam_kind, where.index, _extents(shape), _resolve_dtype(dtype).name, stream, zero)


def empty(shape, dtype="float64", device="cpu", stream=None):
    """An array with uninitialised contents; device memory is allocated stream-ordered on ``stream``."""
    return _allocate(shape, dtype, device, stream, zero=False)


def zeros(shape, dtype="float64", device="cpu", stream=None):
    """An array of zeros (a device fill is queued on ``stream``)."""
    return _allocate(shape, dtype, device, stream, zero=True)


def full(shape, value, dtype="float64", device="cpu", stream=None):
    """An array with every element ``value`` (16-bit floats round to nearest even)."""
    return _allocate(shape, dtype, device, stream, zero=False).fill(value)


def pinned(shape, dtype="float64"):
    """Uninitialised page-locked host memory: copies between it and a device are direct DMA."""
    return _allocate(shape, dtype, "pinned", None, zero=False)


def _template(x, dtype, device):
    """(shape, dtype, device) of a ``*_like`` call: x's own unless overridden; host for non-arrays."""
    here = x.device if isinstance(x, Array) else "cpu"
    return x.shape, (x.dtype if dtype is None else dtype), (here if device is None else device)


def empty_like(x, dtype=None, device=None, stream=None):
    return empty(*_template(x, dtype, device), stream=stream)


def zeros_like(x, dtype=None, device=None, stream=None):
    return zeros(*_template(x, dtype, device), stream=stream)


def from_dlpack(x, stream=None):
    """A view of a foreign array's memory, no copy; the producer's deleter runs once the last view dies.
    ``stream``: the consumer's stream, which the producer orders its pending work before."""
    return _core._from_dlpack(x, stream)


def _adopt(x, stream):
    """x as an Array sharing its memory, or None when nothing about x is viewable. The CUDA array
    interface is tried before DLPack for device objects: it asks no stream ordering, so it is legal inside
    a capture."""
    if isinstance(x, Array):
        return x
    if hasattr(x, "__cuda_array_interface__"):
        return _core._from_cai(x)
    if hasattr(x, "__dlpack_device__"):
        return from_dlpack(x, stream)
    try:
        memoryview(x).release()
    except TypeError:
        return None
    return _core._from_buffer(x)


def _nested_shape(values):
    """The shape of a nested list/tuple, read along its first elements."""
    shape, v = [], values
    while isinstance(v, (list, tuple)):
        shape.append(len(v))
        if not v:
            break
        v = v[0]
    return shape


def _flatten(values, shape):
    """The leaves of a nested list/tuple in row-major order; Error when it is not rectangular."""
    leaves, stack = [], [(values, 0)]
    while stack:
        v, depth = stack.pop()
        if depth == len(shape):
            if isinstance(v, (list, tuple)):
                raise Error("asarray: a ragged nested list has no array shape")
            leaves.append(v)
            continue
        if not isinstance(v, (list, tuple)) or len(v) != shape[depth]:
            raise Error("asarray: a ragged nested list has no array shape")
        stack.extend((e, depth + 1) for e in reversed(v))
    return leaves


def _from_python(values, dtype):
    shape = _nested_shape(values)
    leaves = _flatten(values, shape)
    if dtype is None:  # float if any float, bool if all bool, int otherwise (as numpy infers)
        if any(isinstance(e, float) for e in leaves):
            dtype = "float64"
        elif leaves and all(isinstance(e, bool) for e in leaves):
            dtype = "bool"
        else:
            dtype = "int64"
    out = empty(shape, dtype, "cpu")
    for i, e in enumerate(leaves):
        out._set_flat(i, e)
    return out


def _same_place(a, where):
    here = a.device
    return here.kind == where.kind and (here.kind != "cuda" or here.index == where.index)


def asarray(x, dtype=None, device=None, stream=None):
    """x as an aether array on ``device`` (``None``: wherever it is): no copy when it is already there,
    one copy otherwise. x may be an Array, anything with the buffer protocol, DLPack or the CUDA array
    interface, or a (nested) list of numbers. aether never casts: a dtype other than the data's is an
    Error."""
    src = _adopt(x, stream)
    if src is None:
        src = _from_python(x, dtype)
    if dtype is not None:
        want = _resolve_dtype(dtype)
        if src.dtype != want:
            raise Error(f"asarray: the data is {src.dtype}; aether does no casting (asked for {want})")
    if device is None:
        return src
    where = _place(device)
    if _same_place(src, where):
        return src
    dst = _core._new(where.seam_kind, where.index, list(src.shape), src.dtype.name, stream, False)
    _core._copy(src, dst, stream)
    return dst


def copy(src, into=None, stream=None):
    """Copy ``src`` into ``into`` (a new array on src's device when omitted); the direction follows from
    where the two live. ``stream=None`` copies synchronously, otherwise the copy is queued on ``stream``
    (a copy node inside a capture)."""
    s = _adopt(src, stream)
    if s is None:
        raise Error("copy: src is not an array aether can view")
    if into is None:
        into = empty(s.shape, s.dtype, s.device, stream)
    d = _adopt(into, None)
    if d is None:
        raise Error("copy: into is not an array aether can view")
    _core._copy(s, d, stream)
    return into


from . import fft  # noqa: E402  (the transforms over these arrays)
