"""The `aether-memory/1` table of libaether_hip.so (dsc/src/seam.h:22-51) as a ctypes structure."""
from __future__ import annotations

import ctypes as C

from ._lib import lib

_vp, _i, _sz = C.c_void_p, C.c_int, C.c_size_t

CPU, DEVICE, PINNED = 0, 1, 2
C2C, R2C, C2R = 0, 1, 2
ABI = 2


class Seam(C.Structure):
    _fields_ = [
        ("abi", C.c_uint32),
        ("probe", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("runtime_version", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("current_device", C.CFUNCTYPE(_i, C.POINTER(_i))),
        ("last_error", C.CFUNCTYPE(C.c_char_p)),
        ("chunk_alloc", C.CFUNCTYPE(_i, _i, _i, _sz, _vp, C.POINTER(_vp), C.POINTER(_vp))),
        ("chunk_release", C.CFUNCTYPE(None, _vp)),
        ("drain", C.CFUNCTYPE(None)),
        ("fill", C.CFUNCTYPE(_i, _i, _i, _vp, _sz, _vp, _sz, _vp)),
        ("copy", C.CFUNCTYPE(_i, _i, _i, _vp, _i, _i, _vp, _sz, _vp)),
        ("stream_wait", C.CFUNCTYPE(_i, _i, _vp, _vp)),
        ("stream_sync", C.CFUNCTYPE(_i, _i, _vp)),
        ("pointer_device", C.CFUNCTYPE(_i, _vp, C.POINTER(_i))),
        ("fft_plan", C.CFUNCTYPE(_i, _i, _i, _i, _i, _i, _i, C.POINTER(C.c_int64), _i, C.POINTER(C.c_int32), _vp,
                                 C.POINTER(_vp))),
        ("fft_exec", C.CFUNCTYPE(_i, _vp, _vp, _vp, _vp)),
        ("fft_destroy", C.CFUNCTYPE(None, _vp)),
    ]


_table = None


def table() -> Seam:
    """The library's table, checked like the reference's loader does (dsc/src/aether_core.cpp:95-107)."""
    global _table
    if _table is None:
        fn = lib().aether_memory_v1
        fn.restype = C.POINTER(Seam)
        t = fn().contents
        if t.abi != ABI:
            raise ImportError(f"libaether_hip.so speaks aether-memory/{t.abi}, not /{ABI}")
        _table = t
    return _table


def last_error() -> str:
    msg = table().last_error()
    return msg.decode() if msg else ""
