"""The per-sample members and the elementwise functions against the REFERENCE's own outputs, in double and
in float (tests/golden/ref_members_f64.npz, ref_members_f32.npz, minted from its CPU path by
tests/golden/mint_ref_vectors.py).

The restatements in members_ref.py (which the GPU tests hold the device to, bit for bit) equal the
reference's bits on every arithmetic member; the oracle's transcendental functions are within the reference's
own accuracy of them (the reference's host tanh is up to 2 ulp from the correctly rounded value, the oracle's
within 1; the others within 1 ulp of each other)."""
import hashlib

import numpy as np
import pytest

import members_ref as M

BOUND = {"tanh": 2.0}


@pytest.fixture(scope="module", params=[np.float64, np.float32], ids=["f64", "f32"])
def data(orc, request):
    z = M.load_members(request.param)
    q, r, v = M.members_inputs(orc, z, request.param)
    return z, q, r, v


def test_inputs_regenerate(data):
    z, q, r, v = data
    for name, a in (("q", q), ("r", r), ("v", v)):
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == str(z["sha_" + name])
        assert np.array_equal(a[:, :4], z["head_" + name])


def test_arithmetic_members_equal_reference_bits(data):
    z, q, r, v = data
    for name, want in M.restated(q, r, v).items():
        got = z[name]
        assert got.shape == want.shape, name
        u = np.uint64 if got.dtype == np.float64 else np.uint32
        assert want.dtype == got.dtype, name
        assert np.array_equal(got.view(u), np.ascontiguousarray(want).view(u)), name


@pytest.mark.parametrize("name", M.TRANSCENDENTAL)
def test_transcendental_within_reference_accuracy(orc, data, name):
    z, q, r, v = data
    got = M.transcendental(orc, q, v)[name]
    assert np.max(M.ulps(got, z[name])) <= BOUND.get(name, 1.0), name
