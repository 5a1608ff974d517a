"""Host logic of the transforms' front ends (no GPU): the shared shape rules of aether_amd.dsc._shapes and
the torch front end's norm, dtype and kind rules, as the reference's aether.fft states them
(dsc/aether/fft.py): axes count data axes and wrap, complex data carry a trailing (re, im) axis, r2c keeps
n // 2 + 1 points on the last axis, c2r's real lengths come from s or 2 (m - 1)."""
import pytest
import torch

from aether_amd import fft as F
from aether_amd.dsc import _shapes as S


class Oops(Exception):
    pass


@pytest.mark.parametrize("axes,rank,want", [(None, 3, (0, 1, 2)), (-1, 3, (2,)), ((2, 0), 3, (2, 0)),
                                            ((-3, -2), 3, (0, 1)), (0, 1, (0,))])
def test_data_axes(axes, rank, want):
    assert S.data_axes(axes, rank, Oops) == want


@pytest.mark.parametrize("axes,rank", [(3, 3), (-4, 3), ((0, 5), 2)])
def test_data_axes_out_of_range(axes, rank):
    with pytest.raises(Oops, match="out of range"):
        S.data_axes(axes, rank, Oops)


def test_complex_data_and_real_lengths():
    assert S.complex_data((4, 5, 2), Oops) == (4, 5)
    for bad in ((8,), (4, 3), (2,)):
        with pytest.raises(Oops, match="trailing axis of 2"):
            S.complex_data(bad, Oops)
    assert S.spectrum((4, 9), (1,)) == (4, 5) and S.spectrum((6, 3), (1, 0)) == (4, 3)
    assert S.real_lengths((4, 5), (1,), None, Oops) == (4, 8)
    assert S.real_lengths((4, 5), (1,), (9,), Oops) == (4, 9)
    with pytest.raises(Oops, match="irfft"):
        S.real_lengths((4, 5), (1,), (20,), Oops)
    with pytest.raises(Oops, match="only the last"):
        S.real_lengths((4, 5), (0, 1), (3, 8), Oops)


def test_seam_codes():
    assert [S.seam_code(S.NORMS, n, "norm", Oops) for n in ("backward", "forward", "ortho")] == [0, 1, 2]
    assert [S.seam_code(S.KINDS, k, "kind", Oops) for k in ("c2c", "r2c", "c2r")] == [0, 1, 2]
    for bad in ("bogus", None, 0):
        with pytest.raises(Oops, match="norm"):
            S.seam_code(S.NORMS, bad, "norm", Oops)


def test_torch_front_end_rules():
    assert F._norm("ortho") == 2
    with pytest.raises(F.Error, match="norm"):
        F._norm("unitary")
    assert F._dtype("float32") == torch.float32 and F._dtype(torch.float64) == torch.float64
    with pytest.raises(F.Error, match="float32 or float64"):
        F._dtype("int32")
    with pytest.raises(F.Error, match="kind"):
        F.plan("x2y", "float32", (4,), (0,))
