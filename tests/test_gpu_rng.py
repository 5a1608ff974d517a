"""GPU parity: Philox streams through the C ABI vs the oracle, bit for bit.

Bits and uniforms are integer work plus exact conversions; normals use the
device math, whose algorithm the oracle's own-math route restates operation
for operation, so they are compared bit for bit too (and against the
reference's golden normals within its 1e-12 / 1e-5 band).
"""
import json
import os

import numpy as np
import pytest
import torch

import aether_amd as ae
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def gpu_fill(kind, seed, n, dtype, first=0, offset_elems=0):
    buf = torch.empty(n + offset_elems, dtype=dtype, device="cuda")
    out = buf[offset_elems:]
    ae.rng_fill(kind, seed, out, first=first)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("n,first,off", [(1, 0, 0), (1000, 0, 0), (4097, 3, 0), (123457, 11, 1), (1 << 20, 0, 0)])
def test_bits32(dev, orc, n, first, off):
    got = gpu_fill(ae.RNG_BITS32, 42, n, torch.int32, first, off).view(np.uint32)
    assert np.array_equal(got, orc.fill_bits32(42, n, first))


@pytest.mark.parametrize("kind,dt,ref", [(ae.RNG_UNIFORM_F64, torch.float64, "fill_uniform_f64"),
                                         (ae.RNG_UNIFORM_F32, torch.float32, "fill_uniform_f32"),
                                         (ae.RNG_NORMAL_F64, torch.float64, "fill_normal_f64"),
                                         (ae.RNG_NORMAL_F32, torch.float32, "fill_normal_f32")])
@pytest.mark.parametrize("n,first,off", [(1, 0, 0), (777, 5, 1), (1 << 18, 0, 0)])
def test_streams_bit_exact(dev, orc, kind, dt, ref, n, first, off):
    got = gpu_fill(kind, 7, n, dt, first, off)
    exp = getattr(orc, ref)(7, n, first)
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_golden_rows_on_device(dev):
    rows = json.load(open(os.path.join(GOLDEN, "random_core_rows.json")))
    tol = rows["tolerance"]
    for r in rows["rows"]:
        if r["counter"] != 0:
            continue  # the stream fill draws counter 0; counters are covered by array fills
        g = r["global"]
        bits = gpu_fill(ae.RNG_BITS32, r["seed"], 2, torch.int32, first=2 * g).view(np.uint32)
        assert (int(bits[1]) << 32) | int(bits[0]) == r["bits64"]
        ud = gpu_fill(ae.RNG_UNIFORM_F64, r["seed"], 1, torch.float64, first=g)
        assert int(ud.view(np.uint64)[0]) == r["u01d_bits"]
        nd = gpu_fill(ae.RNG_NORMAL_F64, r["seed"], 1, torch.float64, first=g)[0]
        assert abs(nd - r["normal_d"]) <= tol["normal_d"] * max(1, abs(r["normal_d"]))


@pytest.mark.parametrize("dist", ["uniform", "normal"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_array_fill_counter_is_component(dev, orc, dist, dt):
    n, pitch = 3000, 3008
    arr = ae.Array(n, 3, dt)
    if dist == "uniform":
        ae.array_fill_uniform(arr, 42, -1.0, 1.0)
        exp = orc.fill_array_uniform(42, 3, n, pitch, np.float64 if dt == torch.float64 else np.float32)
    else:
        ae.array_fill_normal(arr, 42, 0.5, 2.0)
        if dt == torch.float32:
            exp = orc.fill_array_normal_f32(42, 3, n, pitch, 0.5, 2.0)
        else:
            exp = orc.fill_array_normal_f64(42, 3, n, pitch, 0.5, 2.0)
    torch.cuda.synchronize()
    got = arr.data.cpu().numpy().reshape(-1)
    exp = exp.reshape(3, pitch)[:, :n]
    assert np.array_equal(got.reshape(3, pitch)[:, :n].view(np.uint8), exp.view(np.uint8))


def test_rejects_out_of_span(dev):
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    with pytest.raises(ae.AetherError):
        ae.rng_fill(ae.RNG_UNIFORM_F64, 1, out, first=(1 << 32) - 2)


def test_generator_normal_leaf_on_device(dev, orc):
    """The Python Generator's normal leaf on the device (E.normal lowered to a tape RNG op) against the
    reference's Generator(42).normal outputs: within the reference's band, and bit for bit the oracle's
    own-math stream (the device's math)."""
    import numpy as np

    from aether_amd import expr as E

    z = np.load(os.path.join(GOLDEN, "ref_generator_normal.npz"))
    n = int(z["n"])
    g = ae.Generator.from_seed(42)
    x = ae.Array(n, 3, torch.float64)
    E.assign(x, E.normal(g, 3, 0.5, 2.0))
    y = ae.Array(n, 3, torch.float32)
    E.assign(y, E.normal(g.substream(3), 3, -1.0, 0.25))
    torch.cuda.synchronize()
    got_d, got_f = x.data[:, :n].cpu().numpy(), y.data[:, :n].cpu().numpy()
    assert np.all(np.abs(got_d - z["d"]) <= 1e-12 * np.maximum(1.0, np.abs(z["d"])))
    assert np.all(np.abs(got_f - z["f"]) <= 1e-5 * np.maximum(1.0, np.abs(z["f"])))
    own_d = orc.fill_array_normal_f64(g.normal_seed(), 3, n, n, 0.5, 2.0, mathmode=1).reshape(3, n)
    own_f = orc.fill_array_normal_f32(g.substream(3).normal_seed(), 3, n, n, -1.0, 0.25, mathmode=1).reshape(3, n)
    assert np.array_equal(got_d.view(np.uint64), own_d.view(np.uint64))
    assert np.array_equal(got_f.view(np.uint32), own_f.view(np.uint32))
