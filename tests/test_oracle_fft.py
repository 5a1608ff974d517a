"""Pins the oracle DFT (long double, tests/test_Fft_common.h:52-77 form).

Closed forms (delta -> constant, constant -> delta, single tone -> one bin),
Parseval, round trip, numpy.fft, and the reference's own pocketfft output
(ref_fft_n*.npz, minted from the reference CPU path) within the reference's
bound (8 + 4 log2 N) eps ||x||_2 (tests/test_Fft_common.h:79-83).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def bound(n, x):
    return (8 + 4 * np.log2(n + 1)) * np.finfo(np.float64).eps * np.linalg.norm(x.view(np.float64))


@pytest.mark.parametrize("n", [1, 2, 8, 64, 1000])
def test_closed_forms(orc, n):
    delta = np.zeros(n, np.complex128)
    delta[0] = 1
    assert np.allclose(orc.dft(delta), np.ones(n), atol=1e-15)
    const = np.ones(n, np.complex128)
    out = orc.dft(const)
    assert abs(out[0] - n) < 1e-12 and np.max(np.abs(out[1:]), initial=0) < 1e-12
    if n > 3:
        k0 = 3
        tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
        out = orc.dft(tone)
        exp = np.zeros(n)
        exp[k0] = n
        assert np.max(np.abs(out - exp)) < 1e-10


@pytest.mark.parametrize("n", [64, 1024])
def test_reference_pocketfft_output(orc, n):
    z = np.load(os.path.join(GOLDEN, f"ref_fft_n{n}.npz"))
    x, y = z["x"], z["y"]
    out = orc.dft(x, forward=True)
    assert np.max(np.abs(out - y)) <= bound(n, x)
    assert np.max(np.abs(np.fft.fft(x) - out)) <= bound(n, x)


def test_parseval_and_round_trip(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(512) + 1j * rng.standard_normal(512)
    X = orc.dft(x)
    assert abs(np.sum(np.abs(X) ** 2) / 512 - np.sum(np.abs(x) ** 2)) < 1e-9
    back = orc.dft(X, forward=False, scale=1 / 512)
    assert np.max(np.abs(back - x)) < 1e-12
