"""Pins the oracle's RNG against the reference's golden data.

Fixtures: tests/golden/random_core_rows.json and random_key_rows.json (the
rows of the reference's tests/random/golden_random.h), ref_rows_*.npz and
ref_keys.npz (minted from the reference's own CPU path), and the published
Random123 known answers for philox4x32-10.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _core():
    return json.load(open(os.path.join(GOLDEN, "random_core_rows.json")))


def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32-10 (Salmon et al., SC'11)
    assert orc.philox((0, 0, 0, 0), 0, 0) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert orc.philox((0xFFFFFFFF,) * 4, 0xFFFFFFFF, 0xFFFFFFFF) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert orc.philox((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), 0xA4093822, 0x299F31D0) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_golden_core_rows_exact(orc):
    # tests/test_Random.cpp:356-388 (Golden_BitsUniformExactBothModes)
    data = _core()
    assert len(data["rows"]) == 315
    for r in data["rows"]:
        s, g, c = r["seed"], r["global"], r["counter"]
        assert orc.bits64(s, g, c) == r["bits64"]
        assert int(np.float64(orc.u01_f64(s, g, c)).view(np.uint64)) == r["u01d_bits"]
        assert int(np.float32(orc.u01_f32(s, g, c)).view(np.uint32)) == r["u01f_bits"]


@pytest.mark.parametrize("mathmode", [0, 1])
def test_golden_normals_within_tolerance(orc, mathmode):
    # tests/test_Random.cpp:390-407: abs AND rel band, 1e-12 (double) / 1e-5 (float)
    data = _core()
    tol = data["tolerance"]
    for r in data["rows"]:
        s, g, c = r["seed"], r["global"], r["counter"]
        nd = orc.normal_f64(s, g, c, mathmode)
        nf = orc.normal_f32(s, g, c, mathmode)
        assert abs(nd - r["normal_d"]) <= tol["normal_d"] * max(1.0, abs(r["normal_d"]))
        assert abs(nf - r["normal_f"]) <= tol["normal_f"] * max(1.0, abs(r["normal_f"]))


def test_golden_key_rows(orc):
    rows = json.load(open(os.path.join(GOLDEN, "random_key_rows.json")))["rows"]
    assert len(rows) == 45
    for r in rows:
        k = orc.generator_key(r["seed"], r["stream"])
        assert k == r["key"]
        assert orc.substream_key(k, r["k"]) == r["sub_key"]


def test_reference_minted_keys(orc):
    rows = np.load(os.path.join(GOLDEN, "ref_keys.npz"))["rows"]
    for seed, stream, k, key, sub in rows.tolist():
        got = orc.generator_key(seed, stream)
        assert got == key
        assert orc.substream_key(got, k) == sub


@pytest.mark.parametrize("seed,counter", [(42, 0), (42, 3), (7, 0), (7, 3)])
def test_reference_minted_rows(orc, seed, counter):
    z = np.load(os.path.join(GOLDEN, f"ref_rows_s{seed}_c{counter}.npz"))
    n = len(z["bits"])
    bits = np.array([orc.bits64(seed, g, counter) for g in range(n)], np.uint64)
    assert np.array_equal(bits, z["bits"])
    ud = np.array([orc.u01_f64(seed, g, counter) for g in range(n)])
    assert np.array_equal(ud.view(np.uint64), z["ud"].view(np.uint64))
    uf = np.array([orc.u01_f32(seed, g, counter) for g in range(n)], np.float32)
    assert np.array_equal(uf.view(np.uint32), z["uf"].view(np.uint32))
    # the libm route reproduces the reference's host normals bit for bit (same glibc)
    nd0 = np.array([orc.normal_f64(seed, g, counter, 0) for g in range(n)])
    assert np.max(np.abs(nd0 - z["nd"]) / np.maximum(1, np.abs(z["nd"]))) <= 1e-15
    nd1 = np.array([orc.normal_f64(seed, g, counter, 1) for g in range(n)])
    assert np.max(np.abs(nd1 - z["nd"]) / np.maximum(1, np.abs(z["nd"]))) <= 1e-12
    nf1 = np.array([orc.normal_f32(seed, g, counter, 1) for g in range(n)], np.float32)
    assert np.max(np.abs(nf1.astype(np.float64) - z["nf"]) / np.maximum(1, np.abs(z["nf"]))) <= 1e-5


def test_streams_and_word_layout(orc):
    # BITS32 word j is word (j & 1) of block(seed, j >> 1, 0): lo then hi of randomBits64
    w = orc.fill_bits32(42, 64)
    for j in range(0, 64, 2):
        b = orc.bits64(42, j // 2, 0)
        assert int(w[j]) == b & 0xFFFFFFFF and int(w[j + 1]) == b >> 32
    u = orc.fill_uniform_f64(42, 16, first=5)
    assert all(u[k] == orc.u01_f64(42, 5 + k, 0) for k in range(16))


BOUNDARY_WORDS = (0, 1, 0x1FF, 0x200, 0x7FFFFFFF, 0x80000000, 0xFFFFFE00, 0xFFFFFFFE, 0xFFFFFFFF)


def test_uniform_open_interval_at_boundary_words(orc):
    """The oracle's word -> real conversions at the reference's boundary words
    (tests/test_RandomIndependence.cpp:98-119): inside (0, 1), monotone in the word, log of the draw and of
    its complement finite; and the known answers at the ends (a wrong shift or offset moves them): double
    2^-33 and 1 - 2^-33, float 2^-24 and 1 - 2^-24, the float keeping the word's top 23 bits."""
    d = [orc.u01_f64_from_word(w) for w in BOUNDARY_WORDS]
    f = [orc.u01_f32_from_word(w) for w in BOUNDARY_WORDS]
    assert all(0.0 < x < 1.0 for x in d) and all(np.float32(0) < x < np.float32(1) for x in f)
    assert d == sorted(d) and f == sorted(f)
    for x, y in zip(d, f):
        assert np.isfinite(np.log(x)) and np.isfinite(np.log(1.0 - x)) and np.isfinite(np.log(np.float32(1) - y))
    assert d[0] == 2.0**-33 and d[-1] == 1.0 - 2.0**-33
    assert f[0] == np.float32(2.0**-24) and f[-1] == np.float32(1.0 - 2.0**-24)
    assert orc.u01_f32_from_word(0x1FF) == f[0] and orc.u01_f32_from_word(0x200) == np.float32(3 * 2.0**-24)
    # the stream's uniform is word 0 of its block, through the same conversion
    for seed, g, c in ((42, 0, 0), (7, 123, 2), (2**64 - 1, 2**32 - 1, 5)):
        b = orc.philox_block(seed, g, c)
        assert orc.u01_f64(seed, g, c) == orc.u01_f64_from_word(b[0])
        assert np.float32(orc.u01_f32(seed, g, c)) == orc.u01_f32_from_word(b[0])


def test_blocks_word_disjoint_across_counters(orc):
    """tests/test_RandomIndependence.cpp:60-96: the 256 blocks of counters 0..255 at one (seed, global) share
    no word, and bits64 is (word 1 << 32) | word 0 of its own block, never overlapping the next counter's."""
    for seed in (0, 0xC0FFEE, 2**64 - 1):
        for g in (0, 3, 2**32 - 1):
            words = [w for c in range(256) for w in orc.philox_block(seed, g, c)]
            assert len(set(words)) == len(words), (seed, g)
        for c in range(0, 512, 7):
            b = orc.philox_block(seed, 7, c)
            assert orc.bits64(seed, 7, c) == (b[1] << 32) | b[0]
            assert orc.bits64(seed, 7, c) >> 32 != orc.bits64(seed, 7, c + 1) & 0xFFFFFFFF


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_normals_independent_across_counters(orc, dt):
    """tests/test_RandomIndependence.cpp:29-57: for independent standard normals E[z^2 z'^2] = 1; the planes
    of an array draw with counter = component, so component 5 against component 5 + lag (lags 1..8, 16, 64)
    over 2^17 samples must give a mean of z^2 z'^2 within 5 standard errors of 1."""
    n, ncomp = 1 << 17, 70
    planes = (orc.fill_array_normal_f64 if dt == "f64" else orc.fill_array_normal_f32)(0x1234567, ncomp, n, n)
    planes = np.asarray(planes, np.float64).reshape(ncomp, n)
    z = planes[5]
    for lag in (1, 2, 3, 4, 5, 6, 7, 8, 16, 64):
        q = z * z * planes[5 + lag] ** 2
        se = np.sqrt(q.var() / n)
        assert abs(q.mean() - 1.0) < 5 * se, (lag, q.mean(), se)


def test_bits_moments(orc):
    # a brute statistical check on a bounded sample (tests/test_Random.cpp moments style)
    u = orc.fill_uniform_f64(42, 1 << 16)
    assert abs(u.mean() - 0.5) < 6 * (1 / np.sqrt(12)) / np.sqrt(u.size)
    z = orc.fill_normal_f64(42, 1 << 16)
    assert abs(z.mean()) < 6 / np.sqrt(z.size)
    assert abs(z.var() - 1) < 6 * np.sqrt(2 / z.size)


def test_generator_normal_leaf_against_the_reference(orc):
    """Generator(42).normal<double,3>(0.5, 2) and .substream(3).normal<float,3>(-1, 0.25), minted from
    the reference's own CPU path (tests/golden/ref_generator_normal.npz): the normal leaf draws with seed
    key ^ 0x13198A2E03707344 (random/detail/DistributionLeaf.h:184), counter = component. The oracle's
    libm route reproduces the reference's bits; its own-math route (the device's) stays within the
    reference's normal band (1e-12 double, 1e-5 float: tests/random/golden_random.h:437-441)."""
    z = np.load(os.path.join(GOLDEN, "ref_generator_normal.npz"))
    n = int(z["n"])
    salt = 0x13198A2E03707344
    key = orc.generator_key(42, 0)
    sub = orc.substream_key(key, 3)
    d0 = orc.fill_array_normal_f64(key ^ salt, 3, n, n, 0.5, 2.0, mathmode=0).reshape(3, n)
    f0 = orc.fill_array_normal_f32(sub ^ salt, 3, n, n, -1.0, 0.25, mathmode=0).reshape(3, n)
    assert np.array_equal(d0.view(np.uint64), z["d"].view(np.uint64))
    assert np.array_equal(f0.view(np.uint32), z["f"].view(np.uint32))
    d1 = orc.fill_array_normal_f64(key ^ salt, 3, n, n, 0.5, 2.0, mathmode=1).reshape(3, n)
    f1 = orc.fill_array_normal_f32(sub ^ salt, 3, n, n, -1.0, 0.25, mathmode=1).reshape(3, n)
    assert np.all(np.abs(d1 - z["d"]) <= 1e-12 * np.maximum(1.0, np.abs(z["d"])))
    assert np.all(np.abs(f1 - z["f"]) <= 1e-5 * np.maximum(1.0, np.abs(z["f"])))
    # the uniform leaf's salt in its place draws a different stream entirely
    wrong = orc.fill_array_normal_f64(key ^ 0x243F6A8885A308D3, 3, n, n, 0.5, 2.0, mathmode=0).reshape(3, n)
    assert np.mean(wrong == z["d"]) < 0.01
