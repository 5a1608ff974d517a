"""GPU: the launch-shape autotuner of the runtime-compiled tape kernels (aeth_tape_tune).

Tuning runs every candidate shape on the caller's arrays; whatever shape it keeps, an assignment without sums
gives the same values as before tuning (bit for bit), the destination of the tuning call holds the result,
and assignments with sums keep the contract shape (their sums equal the specialised kernel's), also after
tuning their occupancy target (aeth_tape_tune_sums). Tuning an
assignment that reads its own destination, or inside a stream capture, is refused by name."""
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu
SHAPES = {(1024, 16, 0), (2048, 16, 0), (4096, 16, 0), (2048, 8, 0), (4096, 8, 0), (2048, 16, 8)}


def filled(n, dt, seeds=(42, 43, 44)):
    out = []
    for s in seeds:
        a = ae.Array(n, 3, dt)
        ae.array_fill_uniform(a, s, -1.0, 1.0)
        out.append(a)
    return out


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [70_001, 1 << 22])
def test_tuned_shape_gives_the_same_values(dev, dt, n):
    a, b, c = filled(n, dt)
    x = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)) + E.tanh(c) * 0.5
    before, tuned, after = (ae.Array(n, 3, dt) for _ in range(3))
    E.assign(before, x)
    got = E.tune(tuned, x)
    assert (got["tile"], got["vector_bytes"], got["waves_per_simd"]) in SHAPES and got["ms"] > 0
    E.assign(after, x)
    torch.cuda.synchronize()
    assert torch.equal(before.data[:, :n], tuned.data[:, :n])
    assert torch.equal(before.data[:, :n], after.data[:, :n])


def test_sums_keep_the_contract_shape_after_tuning(dev):
    n = (1 << 21) + 77
    a, b, c = filled(n, torch.float64, (5, 6, 7))
    x = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))
    d1, d2 = ae.Array(n, 3, torch.float64), ae.Array(n, 3, torch.float64)
    E.tune(d2, x)
    s1 = ae.eval_cfg0(a, b, c, d1, ae.SUM_NEUMAIER)
    s2 = E.assign(d2, x, ae.SUM_NEUMAIER)
    torch.cuda.synchronize()
    assert torch.equal(s1, s2)
    assert torch.equal(d1.data[:, :n], d2.data[:, :n])


@pytest.mark.parametrize("mode", [ae.SUM_NEUMAIER, ae.SUM_PLAIN], ids=["neumaier", "plain"])
def test_tuning_with_sums_keeps_the_bits(dev, mode):
    """With sums only the occupancy target varies: after tuning, the assignment's values and sums equal
    the specialised kernel's (the contract shape), and the pick has the contract tile and width."""
    n = (1 << 22) + 1000
    a, b, c = filled(n, torch.float64, (8, 9, 10))
    x = a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))
    d1, d2 = ae.Array(n, 3, torch.float64), ae.Array(n, 3, torch.float64)
    got = E.tune(d2, x, sum_mode=mode)
    assert (got["tile"], got["vector_bytes"]) == (2048, 16) and got["waves_per_simd"] in (0, 4, 8) and got["ms"] > 0
    s1 = ae.eval_cfg0(a, b, c, d1, mode)
    s2 = E.assign(d2, x, mode)
    torch.cuda.synchronize()
    assert torch.equal(s1, s2)
    assert torch.equal(d1.data[:, :n], d2.data[:, :n])
    with pytest.raises(ae.AetherError, match="destination is also read"):
        E.tune(a, a * 2.0 + b, sum_mode=mode)


def test_tuning_refusals(dev):
    n = 10_000
    a, b, _ = filled(n, torch.float64)
    with pytest.raises(ae.AetherError, match="destination is also read"):
        E.tune(a, a * 2.0 + b)
    d = ae.Array(n, 3, torch.float64)
    E.assign(d, a + b)  # the kernel exists before the capture
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    y = torch.zeros(4, device="cuda")
    with torch.cuda.graph(g, stream=s):
        y += 1
        with pytest.raises(ae.AetherError, match="not allowed inside a stream capture"):
            E.tune(d, a + b, stream=s)
    g.replay()
    torch.cuda.synchronize()
    assert float(y.sum()) == 4.0
