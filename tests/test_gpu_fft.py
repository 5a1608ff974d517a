"""GPU parity: batched uneven FFTs vs the long-double DFT oracle.

Bound (the reference's tests/test_Fft_common.h:79-83):
    max |out - DFT(x) * scale| <= (8 + 4 log2 N) * eps * ||x||_2 * scale
with eps of the storage type. The compensated complex64 modes (complex64 in
memory; float-float or FP64 arithmetic inside) must land far inside the
complex64 bound (stated: within 4 float ulps of |X|max, i.e. the output
rounding), and complex64 plain inside its own bound.
"""
import numpy as np
import pytest
import torch

import aether_amd as ae

pytestmark = pytest.mark.gpu


def bound(n, x, eps, scale=1.0):
    return (8 + 4 * np.log2(n + 1)) * eps * np.linalg.norm(np.asarray(x, np.complex128)) * scale


def run(lengths, x, prec, forward=True, norm=0, offsets=None):
    plan = ae.FftBatch(lengths, offsets, offsets, prec=prec)
    dt = torch.complex128 if prec == 0 else torch.complex64
    xin = torch.from_numpy(x.astype(np.complex128 if prec == 0 else np.complex64)).to("cuda")
    out = torch.zeros_like(xin)
    plan(torch.view_as_real(xin), torch.view_as_real(out), forward, norm)
    torch.cuda.synchronize()
    plan.close()
    assert out.dtype == dt
    return out.cpu().numpy()


def signals(lengths, seed):
    rng = np.random.default_rng(seed)
    tot = int(np.sum(lengths))
    return rng.standard_normal(tot) + 1j * rng.standard_normal(tot)


@pytest.mark.parametrize("log2n", list(range(0, 14)))
def test_single_lengths_c128(dev, orc, log2n):
    n = 1 << log2n
    x = signals([n], log2n)
    for fwd in (True, False):
        y = run([n], x, 0, fwd)
        sc = 1.0 if fwd else 1.0 / n  # BACKWARD norm scales the inverse
        ref = orc.dft(x, forward=fwd, scale=sc)
        assert np.max(np.abs(y - ref)) <= bound(n, x, 2.0**-52, sc)


def test_uneven_batch_c128_norms(dev, orc):
    rng = np.random.default_rng(7)
    lengths = (1 << rng.integers(1, 11, 150)).astype(np.uint64)
    x = signals(lengths, 1)
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    for norm, fwd in ((0, True), (0, False), (1, True), (2, False)):
        y = run(lengths, x, 0, fwd, norm)
        for i in range(0, len(lengths), 7):
            n, o = int(lengths[i]), int(offs[i])
            seg = x[o:o + n]
            sc = 1 / np.sqrt(n) if norm == 2 else (1 / n if fwd == (norm == 1) else 1.0)
            ref = orc.dft(seg, forward=fwd, scale=sc)
            assert np.max(np.abs(y[o:o + n] - ref)) <= bound(n, seg, 2.0**-52, sc)


def test_shuffled_offsets_and_in_place(dev, orc):
    lengths = np.array([64, 8192, 2, 1024, 64, 256, 4096, 8, 64], np.uint64)
    order = np.array([3, 0, 8, 1, 5, 2, 7, 4, 6])
    offs = np.zeros(len(lengths), np.uint64)
    pos = 0
    for i in order:
        offs[i] = pos
        pos += int(lengths[i])
    x = signals(lengths, 3)
    plan = ae.FftBatch(lengths, offs, offs, prec=0)
    buf = torch.from_numpy(x.copy()).to("cuda")
    plan(torch.view_as_real(buf), torch.view_as_real(buf), True, 0)
    torch.cuda.synchronize()
    y = buf.cpu().numpy()
    for i in range(len(lengths)):
        n, o = int(lengths[i]), int(offs[i])
        assert np.max(np.abs(y[o:o + n] - orc.dft(x[o:o + n]))) <= bound(n, x[o:o + n], 2.0**-52)


def test_round_trip_c128(dev):
    lengths = [8192, 4096, 64, 64, 128, 2048]
    x = signals(lengths, 5)
    X = run(lengths, x, 0, True, 0)
    back = run(lengths, X, 0, False, 0)
    assert np.max(np.abs(back - x)) <= 64 * 2.0**-52 * np.max(np.abs(x)) * 13


@pytest.mark.parametrize("log2n", [6, 10, 13])
@pytest.mark.parametrize("prec", [2, 3], ids=["floatfloat", "wide"])
def test_complex64_plain_and_compensated(dev, orc, log2n, prec):
    n = 1 << log2n
    x = signals([n, n], 11).astype(np.complex64)
    ref0 = orc.dft(x[:n].astype(np.complex128))
    y_plain = run([n, n], x, 1)
    y_ff = run([n, n], x, prec)
    err_plain = np.max(np.abs(y_plain[:n] - ref0))
    err_ff = np.max(np.abs(y_ff[:n] - ref0))
    assert err_plain <= bound(n, x[:n], 2.0**-23)
    # compensated: only the final rounding to complex64 remains
    assert err_ff <= 4 * 2.0**-24 * np.max(np.abs(ref0))
    assert err_ff < err_plain


def test_unsupported_length_returns_2(dev):
    with pytest.raises(ae.AetherError) as e:
        ae.FftBatch([12])
    assert e.value.code == 2
