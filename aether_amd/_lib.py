"""ctypes loader for libaether_hip.so (the C ABI declared in include/aether_hip.h).

The product path has no fallback: if the HIP library is missing or fails to
load, every call raises. Pointers are passed as integers (device addresses),
streams as raw hipStream_t handles (0 = null stream).
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libaether_hip.so")

_lock = threading.Lock()
_lib = None

_vp, _u64, _i32, _f64 = C.c_void_p, C.c_uint64, C.c_int, C.c_double

# name -> (restype, argtypes); the single source of truth for the Python side.
SIGNATURES = {
    "aeth_last_error": (C.c_char_p, []),
    "aeth_abi_version": (_i32, []),
    "aeth_probe": (_i32, [C.POINTER(_i32)]),
    "aeth_rng_fill": (_i32, [_i32, _u64, _u64, _u64, _vp, _vp]),
    "aeth_array_fill": (_i32, [_i32, _i32, _u64, _f64, _f64, _i32, _u64, _u64, _vp, _vp]),
    "aeth_eval_cfg0": (_i32, [_i32, _vp, _vp, _vp, _vp, _i32, _u64, _u64, _i32, _vp, _vp]),
    "aeth_eval_cfg0_noalloc": (_i32, [_i32, _vp, _vp, _vp, _vp, _i32, _u64, _u64, _i32, _vp, _vp]),
    "aeth_eval_vec3geo": (_i32, [_vp, _vp, _vp, _vp, _u64, _u64, _vp]),
    "aeth_eval_tape": (_i32, [_i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _u64, _i32, _vp, _vp]),
    "aeth_eval_tape_ex": (_i32, [_i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _u64, _i32, _vp, _vp]),
    "aeth_tape_timing": (_i32, [_i32]),
    "aeth_tape_timing_read": (_i32, [C.POINTER(C.c_float), _i32, C.POINTER(_i32)]),
    "aeth_tape_prepare": (_i32, [_i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _u64, _i32, _vp]),
    "aeth_tape_jit_stats": (_i32, [C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "aeth_tape_jit_build": (_i32, [_i32, _i32, _i32, _vp, C.c_char_p, C.POINTER(_u64)]),
    "aeth_tape_tune": (_i32, [_i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _u64, _vp, _vp]),
    "aeth_tape_tune_sums": (_i32, [_i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _u64, _i32, _vp, _vp, _vp]),
    "aeth_reduce": (_i32, [_i32, _i32, _i32, _vp, _i32, _u64, _u64, _vp, _vp]),
    "aeth_math": (_i32, [_i32, _i32, _vp, _vp, _u64, _vp]),
    "aeth_fft_batch_plan": (_i32, [_i32, _u64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "aeth_fft_batch_exec": (_i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    "aeth_fft_batch_destroy": (_i32, [_vp]),
    "aeth_keyed_sum": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _u64, C.c_int64, _u64, _vp, _vp, _vp, _vp]),
    "aeth_plane_add": (_i32, [_i32, _i32, _vp, _u64, _u64, _vp, _vp, _u64, _vp]),
    "aeth_append_where": (_i32, [_i32, _vp, _u64, _f64, _vp, _u64, _vp, _vp, _i32, _vp]),
    "aeth_runtime_eval": (_i32, [_i32, _vp, _vp, _vp, _f64, _vp]),
}


class AetherError(RuntimeError):
    """A C-ABI entry returned 1 (failure) or 2 (unsupported)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f"libaether_hip.so not built ({LIB_PATH}); run python -m aether_amd.build")
                L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().aeth_last_error().decode(errors="replace")
        raise AetherError(rc, f"{name}: {msg}")
