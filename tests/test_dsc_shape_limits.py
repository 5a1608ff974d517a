"""Shapes whose element or byte count overflows 64 bits are refused by aether_amd.dsc (host side,
no GPU): a wrapped count would allocate a small buffer while the shape, DLPack and the CUDA array
interface report the huge one."""
import numpy as np
import pytest

aether = pytest.importorskip("aether_amd.dsc")


@pytest.mark.parametrize("shape,dtype", [((2**32 + 1, 2**32), "float64"), ((2**31, 2**31, 4), "float32"),
                                         ((2**62, 4), "uint8"), ((2**60,), "float64"), ((2**59, 2), "float64")])
def test_overflowing_shapes_are_refused(shape, dtype):
    with pytest.raises(aether.Error, match="overflows 64 bits"):
        aether.empty(shape, dtype=dtype)
    with pytest.raises(aether.Error, match="overflows 64 bits"):
        aether.zeros(shape, dtype=dtype)


def test_negative_extent_is_refused():
    with pytest.raises(aether.Error, match="negative extent"):
        aether.empty((4, -1), dtype="float64")


def test_largest_representable_shapes_still_describe():
    # zero-sized arrays with huge other extents are fine (no bytes), as are ordinary ones
    a = aether.empty((0, 2**40), dtype="float64")
    assert a.shape == (0, 2**40) and a.size == 0
    b = aether.zeros((3, 5), dtype="float32")
    assert np.asarray(b).shape == (3, 5)
