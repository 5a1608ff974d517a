"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package. Every function
here forwards to the C restatement in oracle/*.c, which cites the reference
file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "aether_ref")

_u64, _u32, _i32, _f64, _f32 = C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_float
_pd = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_pf = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_pu = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

_FAST_PATH = os.path.join(_HERE, "liboracle_fast.so")

_lib = None
_lib_fast = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def lib_fast():
    """The same restatement built with the compiler's default contraction (not bit-exact): the
    performance variant of the CPU baseline only."""
    global _lib_fast
    if _lib_fast is None:
        if not os.path.exists(_FAST_PATH):
            build()
        _lib_fast = _bind(C.CDLL(_FAST_PATH))
    return _lib_fast


def _bind(L):
    sig = {
        "orc_philox4x32_10": (None, [C.POINTER(_u32), _u32, _u32, C.POINTER(_u32)]),
        "orc_philox_block": (None, [_u64, _u32, _u64, C.POINTER(_u32)]),
        "orc_random_bits64": (_u64, [_u64, _u32, _u64]),
        "orc_uniform01_f64": (_f64, [_u64, _u32, _u64]),
        "orc_uniform01_f32": (_f32, [_u64, _u32, _u64]),
        "orc_u01_f64_from_word": (_f64, [_u32]),
        "orc_u01_f32_from_word": (_f32, [_u32]),
        "orc_normal_f64": (_f64, [_u64, _u32, _u64, _i32]),
        "orc_normal_f32": (_f32, [_u64, _u32, _u64, _i32]),
        "orc_splitmix64": (_u64, [_u64]),
        "orc_generator_key": (_u64, [_u64, _u64]),
        "orc_substream_key": (_u64, [_u64, _u64]),
        "orc_fill_bits32": (None, [_u64, _u64, _u64, _pu]),
        "orc_fill_uniform_f64": (None, [_u64, _u64, _u64, _pd]),
        "orc_fill_normal_f64": (None, [_u64, _u64, _u64, _i32, _pd]),
        "orc_fill_uniform_f32": (None, [_u64, _u64, _u64, _pf]),
        "orc_fill_normal_f32": (None, [_u64, _u64, _u64, _i32, _pf]),
        "orc_fill_array_uniform_f64": (None, [_u64, _f64, _f64, _i32, _u64, _u64, _pd]),
        "orc_fill_array_uniform_f32": (None, [_u64, _f32, _f32, _i32, _u64, _u64, _pf]),
        "orc_fill_array_normal_f64": (None, [_u64, _f64, _f64, _i32, _u64, _u64, _i32, _pd]),
        "orc_fill_array_normal_f32": (None, [_u64, _f32, _f32, _i32, _u64, _u64, _i32, _pf]),
        "orc_math_f64": (None, [_i32, _i32, _u64, _pd, _pd]),
        "orc_math_f32": (None, [_i32, _i32, _u64, _pf, _pf]),
        "orc_math2_f64": (None, [_i32, _u64, _pd, _pd, _pd]),
        "orc_math2_f32": (None, [_i32, _u64, _pf, _pf, _pf]),
        "orc_eval_cfg0_f64": (None, [_pd, _pd, _pd, _pd, _i32, _u64, _u64, _i32, _i32]),
        "orc_eval_cfg0_f32": (None, [_pf, _pf, _pf, _pf, _i32, _u64, _u64, _i32, _i32]),
        "orc_eval_vec3geo_f64": (None, [_pd, _pd, _pd, _pd, _u64, _u64, _i32]),
        "orc_tree_sum32_f64": (_f64, [_pd, _u64]),
        "orc_tree_sum32_f32": (_f32, [_pf, _u64]),
        "orc_device_sum_f64": (None, [_pd, _u64, C.POINTER(_f64), C.POINTER(_f64)]),
        "orc_device_sumsq_f64": (None, [_pd, _u64, C.POINTER(_f64), C.POINTER(_f64)]),
        "orc_device_sum_f32": (None, [_pf, _u64, C.POINTER(_f32), C.POINTER(_f32)]),
        "orc_device_sumsq_f32": (None, [_pf, _u64, C.POINTER(_f32), C.POINTER(_f32)]),
        "orc_device_sum_ff_f32": (None, [_pf, _u64, C.POINTER(_f32), C.POINTER(_f32)]),
        "orc_device_sumsq_ff_f32": (None, [_pf, _u64, C.POINTER(_f32), C.POINTER(_f32)]),
        "orc_device_sum_plain_f32": (None, [_pf, _u64, C.POINTER(_f32)]),
        "orc_device_sum_plain_f64": (None, [_pd, _u64, C.POINTER(_f64)]),
        "orc_device_sumsq_plain_f64": (None, [_pd, _u64, C.POINTER(_f64)]),
        "orc_device_sumsq_plain_f32": (None, [_pf, _u64, C.POINTER(_f32)]),
        "orc_neumaier_add_f64": (None, [C.POINTER(_f64), C.POINTER(_f64), _f64]),
        "orc_pair_combine_f64": (None, [_f64, _f64, _f64, _f64, C.POINTER(_f64), C.POINTER(_f64)]),
        "orc_neumaier_add_f32": (None, [C.POINTER(_f32), C.POINTER(_f32), _f32]),
        "orc_pair_combine_f32": (None, [_f32, _f32, _f32, _f32, C.POINTER(_f32), C.POINTER(_f32)]),
        "orc_dft_c128": (None, [_pd, _pd, _u64, _i32, _f64]),
        "orc_block_sum_f64": (None, [_pd, _u64, _i32, C.POINTER(_f64), C.POINTER(_f64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


# ---------------------------------------------------------------- RNG
def philox(ctr, key0, key1):
    c = (_u32 * 4)(*ctr)
    o = (_u32 * 4)()
    lib().orc_philox4x32_10(c, key0, key1, o)
    return tuple(o)


def bits64(seed, g, counter):
    return lib().orc_random_bits64(seed, g, counter)


def u01_f64(seed, g, counter):
    return lib().orc_uniform01_f64(seed, g, counter)


def u01_f32(seed, g, counter):
    return lib().orc_uniform01_f32(seed, g, counter)


def u01_f64_from_word(w):
    return lib().orc_u01_f64_from_word(w)


def u01_f32_from_word(w):
    return np.float32(lib().orc_u01_f32_from_word(w))


def philox_block(seed, g, counter):
    """random/detail/Backend.h:144-150: the block of (seed, global, counter)."""
    o = (_u32 * 4)()
    lib().orc_philox_block(seed, g, counter, o)
    return tuple(o)


def normal_f64(seed, g, counter, mathmode=1):
    return lib().orc_normal_f64(seed, g, counter, mathmode)


def normal_f32(seed, g, counter, mathmode=1):
    return lib().orc_normal_f32(seed, g, counter, mathmode)


def generator_key(seed, stream_id=0):
    return lib().orc_generator_key(seed, stream_id)


def substream_key(key, k):
    return lib().orc_substream_key(key, k)


def fill_bits32(seed, n, first=0):
    out = np.empty(n, np.uint32)
    lib().orc_fill_bits32(seed, first, n, out)
    return out


def fill_uniform_f64(seed, n, first=0):
    out = np.empty(n, np.float64)
    lib().orc_fill_uniform_f64(seed, first, n, out)
    return out


def fill_normal_f64(seed, n, first=0, mathmode=1):
    out = np.empty(n, np.float64)
    lib().orc_fill_normal_f64(seed, first, n, mathmode, out)
    return out


def fill_uniform_f32(seed, n, first=0):
    out = np.empty(n, np.float32)
    lib().orc_fill_uniform_f32(seed, first, n, out)
    return out


def fill_normal_f32(seed, n, first=0, mathmode=1):
    out = np.empty(n, np.float32)
    lib().orc_fill_normal_f32(seed, first, n, mathmode, out)
    return out


def fill_array_uniform(seed, ncomp, n, pitch, dtype=np.float64, lo=-1.0, hi=1.0):
    out = np.zeros(ncomp * pitch, dtype)
    if dtype == np.float64:
        lib().orc_fill_array_uniform_f64(seed, lo, hi, ncomp, n, pitch, out)
    else:
        lib().orc_fill_array_uniform_f32(seed, lo, hi, ncomp, n, pitch, out)
    return out


def fill_array_normal_f64(seed, ncomp, n, pitch, mean=0.0, sd=1.0, mathmode=1):
    out = np.zeros(ncomp * pitch, np.float64)
    lib().orc_fill_array_normal_f64(seed, mean, sd, ncomp, n, pitch, mathmode, out)
    return out


def fill_array_normal_f32(seed, ncomp, n, pitch, mean=0.0, sd=1.0, mathmode=1):
    out = np.zeros(ncomp * pitch, np.float32)
    lib().orc_fill_array_normal_f32(seed, mean, sd, ncomp, n, pitch, mathmode, out)
    return out


# ---------------------------------------------------------------- math
_FN = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "tan": 4, "atan": 5, "asin": 6, "acos": 7, "tanh": 8, "rsqrt": 9,
       "sign": 10}
_FN2 = {"pow": 0, "fmin": 1, "fmax": 2, "selgt": 3}


def math64(fn, x, mode=1):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().orc_math_f64(_FN[fn], mode, x.size, x, y)
    return y


def math2_64(fn, x, y):
    """pow / fmin / fmax / selgt (a > b ? a : b), elementwise over broadcast x, y."""
    x, y = np.broadcast_arrays(np.asarray(x, np.float64), np.asarray(y, np.float64))
    x, y = np.ascontiguousarray(x), np.ascontiguousarray(y)
    z = np.empty_like(x)
    lib().orc_math2_f64(_FN2[fn], x.size, x, y, z)
    return z


def math2_32(fn, x, y):
    x, y = np.broadcast_arrays(np.asarray(x, np.float32), np.asarray(y, np.float32))
    x, y = np.ascontiguousarray(x), np.ascontiguousarray(y)
    z = np.empty_like(x)
    lib().orc_math2_f32(_FN2[fn], x.size, x, y, z)
    return z


def math32(fn, x, mode=1):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_math_f32(_FN[fn], mode, x.size, x, y)
    return y


# ---------------------------------------------------------------- expressions
def eval_cfg0(a, b, c, ncomp, n, pitch, mathmode=1, nthreads=0):
    d = np.zeros_like(a)
    if a.dtype == np.float64:
        lib().orc_eval_cfg0_f64(a, b, c, d, ncomp, n, pitch, mathmode, nthreads)
    else:
        lib().orc_eval_cfg0_f32(a, b, c, d, ncomp, n, pitch, mathmode, nthreads)
    return d


def eval_vec3geo(a, b, c, n, pitch, nthreads=0):
    e = np.zeros(n, np.float64)
    lib().orc_eval_vec3geo_f64(a, b, c, e, n, pitch, nthreads)
    return e


# ---------------------------------------------------------------- sums
def tree_sum32(x):
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return lib().orc_tree_sum32_f64(x, x.size)
    return lib().orc_tree_sum32_f32(x, x.size)


def block_sum(x, nthreads=0, fast=False):
    """Per-thread Neumaier blocks combined in block order (the host's parallel compensated sum)."""
    x = np.ascontiguousarray(x, np.float64)
    h, c = _f64(), _f64()
    (lib_fast() if fast else lib()).orc_block_sum_f64(x, x.size, nthreads, C.byref(h), C.byref(c))
    return h.value, c.value


def device_sum(x, square=False, mode="neumaier"):
    """The device kernels' fixed reduction shape; returns (hi, comp/lo)."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        h, c = _f64(), _f64()
        if mode == "plain":
            (lib().orc_device_sumsq_plain_f64 if square else lib().orc_device_sum_plain_f64)(x, x.size, C.byref(h))
        elif mode == "neumaier":
            (lib().orc_device_sumsq_f64 if square else lib().orc_device_sum_f64)(x, x.size, C.byref(h), C.byref(c))
        else:
            raise ValueError("float-float accumulation is for float32")
        return h.value, c.value
    h, c = _f32(), _f32()
    if mode == "neumaier":
        (lib().orc_device_sumsq_f32 if square else lib().orc_device_sum_f32)(x, x.size, C.byref(h), C.byref(c))
    elif mode == "ff":
        (lib().orc_device_sumsq_ff_f32 if square else lib().orc_device_sum_ff_f32)(x, x.size, C.byref(h),
                                                                                  C.byref(c))
    else:
        (lib().orc_device_sumsq_plain_f32 if square else lib().orc_device_sum_plain_f32)(x, x.size, C.byref(h))
        c.value = 0.0
    return np.float32(h.value), np.float32(c.value)


def neumaier_add(h, c, t, dtype=np.float64):
    if dtype == np.float64:
        H, Cc = _f64(h), _f64(c)
        lib().orc_neumaier_add_f64(C.byref(H), C.byref(Cc), t)
    else:
        H, Cc = _f32(h), _f32(c)
        lib().orc_neumaier_add_f32(C.byref(H), C.byref(Cc), t)
    return H.value, Cc.value


def pair_combine(ah, ac, bh, bc, dtype=np.float64):
    if dtype == np.float64:
        oh, oc = _f64(), _f64()
        lib().orc_pair_combine_f64(ah, ac, bh, bc, C.byref(oh), C.byref(oc))
    else:
        oh, oc = _f32(), _f32()
        lib().orc_pair_combine_f32(ah, ac, bh, bc, C.byref(oh), C.byref(oc))
    return oh.value, oc.value


# ---------------------------------------------------------------- DFT
def dft(x, forward=True, scale=1.0):
    """x: complex128 array; returns complex128 DFT computed in long double."""
    xi = np.ascontiguousarray(np.asarray(x, np.complex128)).view(np.float64)
    out = np.empty_like(xi)
    lib().orc_dft_c128(xi, out, xi.size // 2, 1 if forward else 0, scale)
    return out.view(np.complex128)
