"""Pins the oracle's fused expressions against the reference's own CPU path.

Fixtures ref_cfg0_*.npz and ref_vec3_f64.npz were minted by running the
reference's header-only CPU path (oracle/_ref, tests/golden/mint_ref_vectors.py)
on n = 2000 samples (capacity 2016: pitch != n). With the libm route the
oracle reproduces them bit for bit; the own-math route (the device
algorithm) is within a stated bound; each element is also checked against a
high-precision closed form.
"""
import hashlib
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import GOLDEN

N, PITCH = 2000, 2016


def inputs(orc, seeds, dtype):
    return [orc.fill_array_uniform(s, 3, N, PITCH, dtype) for s in seeds]


def compact(x, dtype):
    return x.reshape(3, PITCH)[:, :N].astype(dtype)


def test_inputs_match_reference(orc):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_f64.npz"))
    a, b, c = inputs(orc, (42, 43, 44), np.float64)
    for arr, key in ((a, "sha_a"), (b, "sha_b"), (c, "sha_c")):
        assert hashlib.sha256(np.ascontiguousarray(compact(arr, np.float64)).tobytes()).hexdigest() == str(z[key])


def test_cfg0_f64_bit_exact_libm_route(orc):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_f64.npz"))
    a, b, c = inputs(orc, (42, 43, 44), np.float64)
    d = orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=0)
    assert np.array_equal(compact(d, np.float64).view(np.uint64), z["d"].view(np.uint64))
    # the reference's canonical compensated sum (32 walks + 32-slot tree)
    for k in range(3):
        assert orc.tree_sum32(compact(d, np.float64)[k]) == z["sums"][k]


def test_cfg0_f32_bit_exact_libm_route(orc):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_f32.npz"))
    a, b, c = inputs(orc, (42, 43, 44), np.float32)
    d = orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=0)
    assert np.array_equal(compact(d, np.float32).view(np.uint32), z["d"].view(np.uint32))


def test_cfg0_generator_inputs(orc):
    # Generator(42).uniform for a; substream(1), substream(2) for b and c (random/Generator.h:64-90)
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_generator_f64.npz"))
    k = orc.generator_key(42, 0)
    seeds = [k ^ 0x243F6A8885A308D3, orc.substream_key(k, 1) ^ 0x243F6A8885A308D3,
             orc.substream_key(k, 2) ^ 0x243F6A8885A308D3]
    a, b, c = inputs(orc, seeds, np.float64)
    d = orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=0)
    assert np.array_equal(compact(d, np.float64).view(np.uint64), z["d"].view(np.uint64))


def test_cfg0_own_math_close_and_closed_form(orc):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_f64.npz"))
    a, b, c = inputs(orc, (42, 43, 44), np.float64)
    d = compact(orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=1), np.float64)
    ref = z["d"]
    # bound: sin and exp each within 1 ulp, |terms| <= 2.72 -> a few ulps of the result scale
    scale = np.abs(compact(a, np.float64)) + np.abs(compact(b * c, np.float64)) + 1.0
    assert np.max(np.abs(d - ref) / (scale * 2.0**-52)) <= 4.0
    A, B, C = (compact(x, np.float64) for x in (a, b, c))
    with mp.workprec(120):
        for k, i in [(0, 0), (1, 17), (2, 1999), (0, 1000), (2, 77)]:
            exact = mp.mpf(A[k, i]) + mp.mpf(B[k, i]) * mp.mpf(C[k, i]) + mp.sin(A[k, i]) * mp.exp(-mp.mpf(B[k, i])**2)
            assert abs(float(exact) - d[k, i]) <= 4 * 2.0**-52 * scale[k, i]


def test_vec3_bit_exact(orc):
    z = np.load(os.path.join(GOLDEN, "ref_vec3_f64.npz"))
    a, b, c = inputs(orc, (42, 43, 44), np.float64)
    e = orc.eval_vec3geo(a, b, c, N, PITCH)
    assert np.array_equal(e.view(np.uint64), z["e"].view(np.uint64))
    A, B, C = (compact(x, np.float64) for x in (a, b, c))
    closed = np.einsum("ki,ki->i", np.cross(A.T, B.T).T, C) + np.sqrt((A * A).sum(0) + 1)
    assert np.max(np.abs(closed - e)) <= 1e-14


@pytest.mark.parametrize("nthreads", [1, 3])
def test_eval_thread_invariant(orc, nthreads):
    a, b, c = inputs(orc, (42, 43, 44), np.float64)
    d1 = orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=1, nthreads=nthreads)
    d2 = orc.eval_cfg0(a, b, c, 3, N, PITCH, mathmode=1, nthreads=8)
    assert np.array_equal(d1, d2)
