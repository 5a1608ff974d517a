"""The torch-facing API of aether_amd (re-exported lazily by the package).

Host-side mirror of the reference interface over the C ABI of
libaether_hip.so (include/aether_hip.h). torch is used only to own device
memory and streams; every computation runs in the hand-written HIP kernels.

Layout: an ``Array`` of ``ncomp`` components and ``n`` samples is a device
buffer of ``ncomp`` planes of ``capacity`` elements, capacity = n rounded up
to 32 (array/Array.h:80,100-110,279-295); element (c, i) is at c*capacity + i.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from ._lib import AetherError, call, lib

__all__ = [
    "AetherError", "Array", "Generator", "rng_fill", "array_fill_uniform", "array_fill_normal", "eval_cfg0",
    "eval_vec3geo", "reduce", "math", "stream_handle", "F64", "F32", "SUM_NONE", "SUM_NEUMAIER", "SUM_PLAIN",
    "SUM_FLOATFLOAT", "FftBatch", "RNG_BITS32", "RNG_UNIFORM_F64", "RNG_UNIFORM_F32", "RNG_NORMAL_F64", "RNG_NORMAL_F32",
]

F64, F32 = 0, 1
SUM_NONE, SUM_NEUMAIER, SUM_PLAIN, SUM_FLOATFLOAT = -1, 0, 1, 2
RNG_BITS32, RNG_UNIFORM_F64, RNG_UNIFORM_F32, RNG_NORMAL_F64, RNG_NORMAL_F32 = 0, 1, 2, 3, 4
AETH_RNG_DIST_UNIFORM, AETH_RNG_DIST_NORMAL = 0, 1
FN = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "all4": 4}

CAPACITY_ALIGNMENT = 32  # array/Array.h:80 kCapacityAlignment

# Per-distribution salts (random/detail/DistributionLeaf.h:117,184) and the
# substream salt (random/Generator.h:193).
UNIFORM_SALT = 0x243F6A8885A308D3
NORMAL_SALT = 0x13198A2E03707344
SUBSTREAM_SALT = 0x2545F4914F6CDD1D
DEFAULT_SEED = 0xFE7A5EED0C0FFEE5  # random/free.h:44
_M64 = (1 << 64) - 1


def _dtype_code(t: torch.dtype) -> int:
    if t == torch.float64:
        return F64
    if t == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {t}")


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def splitmix64(x: int) -> int:
    """random/detail/Backend.h:75-81 (host arithmetic for keys only)."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


@dataclass(frozen=True)
class Generator:
    """random::Generator (random/Generator.h:56-195): a 64-bit key."""

    key: int

    @staticmethod
    def from_seed(seed: int, stream_id: int = 0) -> "Generator":
        return Generator(splitmix64((seed ^ splitmix64(stream_id)) & _M64))

    def substream(self, k: int) -> "Generator":
        return Generator(splitmix64(self.key ^ splitmix64(k ^ SUBSTREAM_SALT)))

    def uniform_seed(self) -> int:
        return self.key ^ UNIFORM_SALT

    def normal_seed(self) -> int:
        return self.key ^ NORMAL_SALT


class Array:
    """Array<T, ncomp> resident on the device (array/Array.h:82-364)."""

    def __init__(self, n: int, ncomp: int = 3, dtype: torch.dtype = torch.float64, device="cuda"):
        if n < 0 or ncomp < 1:
            raise ValueError("bad shape")
        self.n = int(n)
        self.ncomp = int(ncomp)
        self.capacity = -(-self.n // CAPACITY_ALIGNMENT) * CAPACITY_ALIGNMENT
        self.data = torch.empty((self.ncomp, max(self.capacity, 1)), dtype=dtype, device=device)

    @property
    def dtype(self):
        return self.data.dtype

    def samples(self) -> int:
        return self.n

    def ptr(self) -> int:
        return self.data.data_ptr()

    def plane(self, c: int) -> torch.Tensor:
        return self.data[c, : self.n]

    def host(self):
        """(ncomp, n) host copy (download, array/Array.h:258-268)."""
        return self.data[:, : self.n].cpu()

    # an Array in an expression is a leaf (expr/Operations.h:25-71)
    def _leaf(self):
        from .expr import Leaf

        return Leaf(self)

    def __add__(self, o): return self._leaf() + o
    def __radd__(self, o): return o + self._leaf()
    def __sub__(self, o): return self._leaf() - o
    def __rsub__(self, o): return o - self._leaf()
    def __mul__(self, o): return self._leaf() * o
    def __rmul__(self, o): return o * self._leaf()
    def __truediv__(self, o): return self._leaf() / o
    def __neg__(self): return -self._leaf()
    def dot(self, o): return self._leaf().dot(o)
    def cross(self, o): return self._leaf().cross(o)


def rng_fill(kind: int, seed: int, out: torch.Tensor, first: int = 0, stream=None) -> torch.Tensor:
    call("aeth_rng_fill", kind, seed & _M64, first, out.numel(), out.data_ptr(), stream_handle(stream))
    return out


def array_fill_uniform(arr: Array, seed: int, lo=-1.0, hi=1.0, stream=None) -> Array:
    call("aeth_array_fill", 0, _dtype_code(arr.dtype), seed & _M64, float(lo), float(hi), arr.ncomp, arr.n,
         arr.capacity, arr.ptr(), stream_handle(stream))
    return arr


def array_fill_normal(arr: Array, seed: int, mean=0.0, sd=1.0, stream=None) -> Array:
    call("aeth_array_fill", 1, _dtype_code(arr.dtype), seed & _M64, float(mean), float(sd), arr.ncomp, arr.n,
         arr.capacity, arr.ptr(), stream_handle(stream))
    return arr


def eval_cfg0(a: Array, b: Array, c: Array, d: Array, sum_mode: int = SUM_NEUMAIER, sums: torch.Tensor = None,
              stream=None):
    """d = a + b*c + sin(a)*exp(-b*b); returns the (ncomp, 2) (hi, comp) sums tensor or None."""
    for x in (b, c, d):
        if (x.n, x.ncomp, x.capacity, x.dtype) != (a.n, a.ncomp, a.capacity, a.dtype):
            raise ValueError("operand shapes differ")
    if sum_mode != SUM_NONE and sums is None:
        sums = torch.empty((a.ncomp, 2), dtype=a.dtype, device=a.data.device)
    call("aeth_eval_cfg0", _dtype_code(a.dtype), a.ptr(), b.ptr(), c.ptr(), d.ptr(), a.ncomp, a.n, a.capacity,
         sum_mode, sums.data_ptr() if sums is not None else None, stream_handle(stream))
    return sums


def eval_vec3geo(a: Array, b: Array, c: Array, e: torch.Tensor, stream=None) -> torch.Tensor:
    if a.dtype != torch.float64 or a.ncomp != 3:
        raise TypeError("vec3geo takes Array<double,3>")
    call("aeth_eval_vec3geo", a.ptr(), b.ptr(), c.ptr(), e.data_ptr(), a.n, a.capacity, stream_handle(stream))
    return e


def reduce(x: Array, mode: int = SUM_NEUMAIER, square: bool = False, out: torch.Tensor = None, stream=None):
    if out is None:
        out = torch.empty((x.ncomp, 2), dtype=x.dtype, device=x.data.device)
    call("aeth_reduce", _dtype_code(x.dtype), mode, int(square), x.ptr(), x.ncomp, x.n, x.capacity,
         out.data_ptr(), stream_handle(stream))
    return out


def math(fn: str, x: torch.Tensor, y: torch.Tensor = None, stream=None) -> torch.Tensor:
    n = x.numel()
    if y is None:
        y = torch.empty((4 * n,) if fn == "all4" else (n,), dtype=x.dtype, device=x.device)
    call("aeth_math", FN[fn], _dtype_code(x.dtype), x.data_ptr(), y.data_ptr(), n, stream_handle(stream))
    return y


class FftBatch:
    """A batch of uneven power-of-two complex FFTs run as one launch (aeth_fft_batch_*).

    prec: 0 complex128, 1 complex64, 2 complex64 with float-float arithmetic (the compensated FP32 mode), 3 complex64 in memory with FP64
    arithmetic.
    lengths / offsets are in complex elements of the flat input/output buffers.
    """

    def __init__(self, lengths, in_offsets=None, out_offsets=None, prec: int = 0):
        import ctypes as C

        import numpy as np

        L = np.ascontiguousarray(lengths, np.uint64)
        if in_offsets is None:
            in_offsets = np.concatenate([[0], np.cumsum(L)[:-1]]).astype(np.uint64) if len(L) else L
        I = np.ascontiguousarray(in_offsets, np.uint64)
        O = np.ascontiguousarray(out_offsets if out_offsets is not None else I, np.uint64)
        self.prec = prec
        self.total = int(L.sum())
        h = C.c_void_p()
        call("aeth_fft_batch_plan", prec, len(L), L.ctypes.data, I.ctypes.data, O.ctypes.data, C.byref(h))
        self._h = h

    def __call__(self, x: torch.Tensor, out: torch.Tensor, forward: bool = True, norm: int = 0, stream=None):
        call("aeth_fft_batch_exec", self._h, x.data_ptr(), out.data_ptr(), int(forward), norm, stream_handle(stream))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            call("aeth_fft_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
