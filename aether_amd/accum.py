"""Accumulation over device arrays: keyed sums with one atomic per run of lanes (accum/run.h),
compensated accumulation planes (accum/AccumPlane.h) and append counters (accum/AppendCounter.h).

The device primitives are HIP code in include/aether/device/accum.hpp (wave64 forms); these wrappers run
them over whole arrays through the C ABI (aeth_keyed_sum, aeth_plane_add, aeth_append_where).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import F32, F64, Array, _seam, stream_handle
from ._lib import AetherError, call

__all__ = ["keyed_sum", "Plane", "DeviceFlag", "AppendCounter", "append_where", "ALIGNED", "TESTED", "SORTED"]

ALIGNED, TESTED, SORTED = 0, 1, 2
_FORMS = {"aligned": ALIGNED, "tested": TESTED, "sorted": SORTED}


def _code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise AetherError(1, f"terms are float32 or float64, not {t.dtype}")


def _fill_zero(ptr: int, nbytes: int, stream=None) -> None:
    """Zero device memory with the seam's fill kernel (capture-safe, no memset node)."""
    if nbytes == 0:
        return
    zero = C.c_uint32(0)
    t = _seam.table()
    rc = t.fill(_seam.DEVICE, torch.cuda.current_device(), C.c_void_p(ptr), nbytes, C.byref(zero), 4,
                C.c_void_p(stream_handle(stream)))
    if rc != 0:
        raise AetherError(rc, _seam.last_error())


def keyed_sum(keys: torch.Tensor, terms: torch.Tensor, slots: int, form="sorted", width: int = 64,
              compensated: bool = False, own: int = 0, target: torch.Tensor = None, comp: torch.Tensor = None,
              count_commits: bool = False, stream=None):
    """target[keys[l]] += terms[l] (plus the residuals into comp when compensated), one atomic per run.

    Returns (target, comp, commits): commits is a (run, lane) pair of commit counts when count_commits,
    else None. target / comp are zeroed new arrays unless given (then they accumulate)."""
    f = _FORMS[form] if isinstance(form, str) else int(form)
    if keys.dtype != torch.int64 or not keys.is_cuda or not terms.is_cuda:
        raise AetherError(1, "keys are int64 and keys / terms live on the GPU")
    keys, terms = keys.contiguous(), terms.contiguous()
    n = terms.numel()
    if keys.numel() != n:
        raise AetherError(1, "one key per term")
    if target is None:
        target = torch.empty(slots, dtype=terms.dtype, device=terms.device)
        _fill_zero(target.data_ptr(), target.numel() * target.element_size(), stream)
    if compensated and comp is None:
        comp = torch.empty(slots, dtype=terms.dtype, device=terms.device)
        _fill_zero(comp.data_ptr(), comp.numel() * comp.element_size(), stream)
    commits = None
    if count_commits:
        commits = torch.empty(2, dtype=torch.int64, device=terms.device)
        _fill_zero(commits.data_ptr(), 16, stream)
    call("aeth_keyed_sum", _code(terms), f, width, int(bool(compensated)), keys.data_ptr(), terms.data_ptr(), n,
         own, slots, target.data_ptr(), comp.data_ptr() if comp is not None else None,
         commits.data_ptr() if commits is not None else None, stream_handle(stream))
    return target, comp, commits


class Plane:
    """AccumPlane<double, vd> (accum/AccumPlane.h:405-470): value and companion planes of `n` samples."""

    def __init__(self, n: int, vd: int = 3, policy: str = "compensated"):
        if policy not in ("compensated", "serialized"):
            raise AetherError(1, f"policy is 'compensated' or 'serialized', not {policy!r}")
        self.n, self.vd, self.policy = n, vd, policy
        self.value = Array(n, vd, torch.float64)
        self.comp = Array(n if policy == "compensated" else 1, vd, torch.float64)
        self.zero()

    def zero(self, stream=None) -> "Plane":
        """Both lanes, unconditionally (zeroDeviceAsync: accum/AccumPlane.h:448-460)."""
        for a in (self.value, self.comp):
            _fill_zero(a.ptr(), a.data.numel() * 8, stream)
        return self

    def add(self, terms: Array, stream=None) -> "Plane":
        """plane[i] += terms[:, i] for every sample (components in ascending order)."""
        if terms.ncomp != self.vd or terms.n != self.n or terms.dtype != torch.float64:
            raise AetherError(1, "terms must be an Array<double, vd> of the plane's samples")
        call("aeth_plane_add", 0 if self.policy == "compensated" else 1, self.vd, terms.ptr(), terms.capacity, self.n,
             self.value.ptr(), self.comp.ptr() if self.policy == "compensated" else None, self.value.capacity,
             stream_handle(stream))
        return self

    def delivered(self, out: Array = None, stream=None) -> Array:
        """value + comp per sample and lane (one fused kernel)."""
        from . import expr as E

        if out is None:
            out = Array(self.n, self.vd, torch.float64)
        if self.policy == "compensated":
            E.assign(out, self.value + self.comp, stream=stream)
        else:
            E.assign(out, self.value * 1.0, stream=stream)
        return out


class DeviceFlag:
    """The sticky device error word (err/DeviceFlag.h:86-160): kernels set it (an atomic OR of 1 through
    `ptr`, DeviceFlagView in include/aether/device/flag.hpp), the host reads it (`is_set`, never clearing
    it), clears it, or turns it into an error (`check_and_clear`)."""

    def __init__(self, device="cuda"):
        self._mem = torch.empty(4, dtype=torch.int32, device=device)  # one word on a 16-byte slot
        self.clear()

    @property
    def ptr(self) -> int:
        return self._mem.data_ptr()

    def is_set(self) -> bool:
        torch.cuda.synchronize(self._mem.device)
        return bool(self._mem[0].item())

    def clear(self) -> None:
        torch.cuda.synchronize(self._mem.device)
        _fill_zero(self._mem.data_ptr(), 16)
        torch.cuda.synchronize(self._mem.device)

    def check_and_clear(self, context: str) -> None:
        """If set: clear it and raise AetherError shaped like err::fail (err/Error.h:63)."""
        if self.is_set():
            self.clear()
            raise AetherError(1, f"{context} failed on DeviceFlag, size=4 bytes (device flag was set)")


class AppendCounter:
    """A device slot counter over `capacity` places with a DeviceFlag for overflow
    (accum/AppendCounter.h:82-192)."""

    def __init__(self, capacity: int, device="cuda"):
        self.capacity = int(capacity)
        self._mem = torch.empty(4, dtype=torch.int32, device=device)  # the counter on a 16-byte slot
        self.overflow = DeviceFlag(device)
        self.reset()

    @property
    def counter_ptr(self) -> int:
        return self._mem.data_ptr()

    @property
    def overflow_ptr(self) -> int:
        return self.overflow.ptr

    def raw_count(self) -> int:
        """The raw counter (may exceed capacity: every overflowing claim still incremented it)."""
        return int(self._mem[0].item()) & 0xFFFFFFFF

    def overflowed(self) -> bool:
        """Non-destructive peek at the overflow flag."""
        return self.overflow.is_set()

    def reset(self, clear_overflow: bool = False, stream=None) -> None:
        _fill_zero(self._mem.data_ptr(), 16, stream)
        if clear_overflow:
            self.overflow.clear()

    def check_overflow(self, context: str = "AppendCounter::checkOverflow") -> None:
        """Raise if the overflow flag is set, and clear it (AppendCounter::checkOverflow)."""
        self.overflow.check_and_clear(context)

    def commit(self) -> int:
        """The claimed count clamped to capacity; resets the counter for the next step (never raises)."""
        k = min(self.raw_count(), self.capacity)
        self.reset()
        return k


CLAIMS = {"lane": 0, "wave": 1, "tile": 2}


def append_where(x: torch.Tensor, threshold: float, counter: AppendCounter, out: torch.Tensor,
                 claim: str = "tile", stream=None) -> None:
    """Append the index of every element of x above `threshold` to `out` (int32, counter.capacity places).
    claim: one atomic per "lane" (claim), per "wave" or per "tile" of 16384 elements."""
    if out.dtype != torch.int32 or out.numel() < counter.capacity or not out.is_cuda:
        raise AetherError(1, "out is an int32 device tensor with the counter's capacity")
    x = x.contiguous()
    call("aeth_append_where", _code(x), x.data_ptr(), x.numel(), float(threshold), out.data_ptr(), counter.capacity,
         counter.counter_ptr, counter.overflow_ptr, CLAIMS[claim], stream_handle(stream))
