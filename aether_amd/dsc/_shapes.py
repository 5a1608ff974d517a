"""Argument and shape rules of the transforms, shared by ``aether_amd.dsc.fft`` (aether arrays) and
``aether_amd.fft`` (torch tensors); plain Python. Each rule raises ``error(message)``, the caller's own
exception type.

Complex data carry a trailing (re, im) axis; axes count the other ("data") axes. A real-to-complex
transform keeps n // 2 + 1 points along the last transformed axis, and its inverse takes the real lengths
from ``s`` or, by default, 2 (m - 1) on that axis (numpy's conventions, which the reference's
dsc/aether/fft.py follows)."""

KINDS = ("c2c", "r2c", "c2r")  # seam codes by position (dsc/src/seam.h:18)
NORMS = ("backward", "forward", "ortho")  # seam codes by position (dsc/src/seam.h:19)


def seam_code(table, value, what, error):
    if isinstance(value, str) and value in table:
        return table.index(value)
    raise error(f"{what}= must be one of {', '.join(repr(t) for t in table)}; got {value!r}")


def data_axes(axes, rank, error):
    """Axes over ``rank`` data axes (None: all of them), each wrapped into [0, rank)."""
    if axes is None:
        return tuple(range(rank))
    picked = (axes,) if isinstance(axes, int) else tuple(axes)
    for a in picked:
        if a < -rank or a >= rank:
            raise error(f"axis {a} is out of range: the data have {rank} axes")
    return tuple(a % rank for a in picked)


def complex_data(shape, error):
    """The data shape of complex data: every axis but the trailing (re, im) one."""
    dims = tuple(shape)
    if len(dims) < 2 or dims[-1] != 2:
        raise error(f"complex data need a trailing axis of 2 (re, im); shape {dims} has none")
    return dims[:-1]


def spectrum(dims, axes):
    """The complex side's data shape of a real transform of ``dims`` over ``axes``."""
    last = axes[-1]
    return dims[:last] + (dims[last] // 2 + 1,) + dims[last + 1:]


def real_lengths(dims, axes, s, error):
    """The real side's data shape of a complex-to-real transform of complex data ``dims``."""
    real = list(dims)
    if s is None:
        real[axes[-1]] = 2 * (dims[axes[-1]] - 1)
    else:
        for a, length in zip(axes, s):
            real[a] = int(length)
    last = real[axes[-1]]
    if last // 2 + 1 != dims[axes[-1]]:
        raise error(f"irfft: {dims[axes[-1]]} complex points cannot come from a real length of {last} "
                    f"(which has {last // 2 + 1})")
    for a in axes[:-1]:
        if real[a] != dims[a]:
            raise error("irfftn: only the last of the axes changes length (the others are complex-to-complex)")
    return tuple(real)
