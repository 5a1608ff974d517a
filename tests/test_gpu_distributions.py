"""GPU: Generator's lognormal, exponential and bernoulli draws (random/Generator.h:113-165,
random/detail/DistributionLeaf.h:240-430) lowered to the tape, bit for bit against the oracle's draws of the
same streams (each distribution's own salt, counter = component) passed through the oracle's math functions,
in double and in float. A bernoulli plane holds only 0 and 1, at the expected rate."""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def same_bits(got, want):
    u = np.uint64 if got.dtype == np.float64 else np.uint32
    return bool(np.all((got.view(u) == want.view(u)) | (np.isnan(got) & np.isnan(want))))


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_distributions_bit_exact(dev, orc, dt):
    n, nc = 100_003, 3
    g = ae.Generator.from_seed(2024).substream(3)
    f32 = dt == torch.float32
    ft = np.float32 if f32 else np.float64
    m1 = orc.math32 if f32 else orc.math64

    def uniform(salt):
        return np.asarray(orc.fill_array_uniform(g.key ^ salt, nc, n, n, ft, 0.0, 1.0)).reshape(nc, n)

    def normal(salt, mu, sigma):
        fill = orc.fill_array_normal_f32 if f32 else orc.fill_array_normal_f64
        return np.asarray(fill(g.key ^ salt, nc, n, n, mu, sigma)).reshape(nc, n)

    def fn(name, x):
        return m1(name, np.ascontiguousarray(x.ravel())).reshape(x.shape)

    u_exp, u_ber = uniform(E.EXPONENTIAL_SALT), uniform(E.BERNOULLI_SALT)
    cases = {
        "lognormal": (E.lognormal(g, nc, 0.25, 0.5), fn("exp", normal(E.LOGNORMAL_SALT, 0.25, 0.5))),
        "exponential": (E.exponential(g, nc, 1.5), -fn("log", u_exp) / ft(1.5)),
        "bernoulli": (E.bernoulli(g, nc, 0.3), (u_ber < ft(0.3)).astype(ft)),
    }
    for name, (tree, want) in cases.items():
        d = ae.Array(n, nc, dt)
        E.assign(d, tree)
        got = d.data[:, :n].cpu().numpy()
        assert same_bits(got, want.astype(ft)), name
    b = cases["bernoulli"][1]
    assert set(np.unique(b)) <= {0.0, 1.0} and abs(b.mean() - 0.3) < 6 * np.sqrt(0.3 * 0.7 / b.size)
