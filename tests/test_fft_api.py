"""Host logic of aether_amd.fft (no GPU): axis, norm and dtype rules as dsc/aether/fft.py states them."""
import pytest
import torch

from aether_amd import fft as F


def test_axes_rules():
    assert F._axes(None, 3) == (0, 1, 2)
    assert F._axes(-1, 3) == (2,)
    assert F._axes((2, 0), 3) == (2, 0)
    with pytest.raises(F.Error, match="out of range"):
        F._axes(3, 3)


def test_norm_and_dtype_rules():
    assert [F._norm(n) for n in ("backward", "forward", "ortho")] == [0, 1, 2]
    with pytest.raises(F.Error, match="norm"):
        F._norm("bogus")
    assert F._dtype("float32") == torch.float32 and F._dtype(torch.float64) == torch.float64
    with pytest.raises(F.Error, match="float32 or float64"):
        F._dtype("int32")
    with pytest.raises(F.Error, match="kind"):
        F.plan("x2y", "float32", (4,), (0,))
