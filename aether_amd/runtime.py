"""The runtime elementwise evaluator over strided N-d views (eval/Runtime.h, view/RuntimeView.h,
interop/DLPack.h): ``runtime_eval(op, a, b, out, scalar)`` with ops assign (the one that casts
double <-> float), add, sub, scale, add_scaled, sub_scaled. Operands are device tensors of one logical
shape, each with its own strides (transposes, slices, broadcasts by stride 0); anything with
``__dlpack__`` is taken through DLPack without a copy.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import F32, F64, stream_handle
from ._lib import AetherError, call

__all__ = ["RuntimeView", "view", "runtime_eval", "OPS", "RANK_CAP"]

RANK_CAP = 6
OPS = {"assign": 0, "add": 1, "sub": 2, "scale": 3, "add_scaled": 4, "sub_scaled": 5}


class RuntimeView(C.Structure):
    """aeth_runtime_view: data, dtype, rank, extents and element strides (view/RuntimeView.h:77-93)."""

    _fields_ = [("data", C.c_void_p), ("dtype", C.c_int32), ("rank", C.c_int32),
                ("extents", C.c_uint32 * RANK_CAP), ("strides", C.c_int64 * RANK_CAP)]


def _tensor(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__"):
        return torch.from_dlpack(x)
    raise AetherError(1, f"not a device array: {type(x)}")


def view(x) -> RuntimeView:
    """A RuntimeView of a device tensor (or DLPack producer), without a copy."""
    t = _tensor(x)
    if not t.is_cuda:
        raise AetherError(1, "runtime views are of device memory")
    if t.dim() > RANK_CAP:
        raise AetherError(1, f"rank {t.dim()} exceeds the cap of {RANK_CAP}")
    code = {torch.float64: F64, torch.float32: F32}.get(t.dtype)
    if code is None:
        raise AetherError(1, f"unsupported dtype {t.dtype} (double/float only)")
    v = RuntimeView()
    v.data = t.data_ptr()
    v.dtype = code
    v.rank = t.dim()
    for k, (e, st) in enumerate(zip(t.shape, t.stride())):
        v.extents[k] = e
        v.strides[k] = st
    v._keep = t  # the view borrows the tensor's memory
    return v


def runtime_eval(op: str, a, b=None, out=None, scalar: float = 0.0, stream=None):
    """out = op(a, b, scalar) elementwise; returns out (a new dense tensor of a's shape and dtype when None)."""
    if op not in OPS:
        raise AetherError(1, f"op is one of {sorted(OPS)}, not {op!r}")
    ta = _tensor(a)
    if out is None:
        out = torch.empty(ta.shape, dtype=ta.dtype, device=ta.device)
    va, vo = view(ta), view(out)
    vb = view(b) if b is not None else None
    call("aeth_runtime_eval", OPS[op], C.byref(va), C.byref(vb) if vb is not None else None, C.byref(vo),
         float(scalar), stream_handle(stream))
    return out
