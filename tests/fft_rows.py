"""Shared pieces of the FFT front-end tests: seeded data, conversions between (re, im) pairs and numpy
complex, and the reference's error bound for a transform over N points,
max |error| <= (8 + 4 log2(N + 1)) * eps * ||x||_2 * s  (dsc/tests/test_fft.py states it), with s the
transform's scale: 1, 1/N or 1/sqrt(N) by norm and direction."""
import math

import numpy as np

EPS = {"float32": float(np.finfo(np.float32).eps), "float64": float(np.finfo(np.float64).eps)}
LENGTHS = (1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 17, 31, 64, 100, 127, 1000)
REAL_LENGTHS = (2, 3, 7, 8, 9, 16, 17, 50, 128)
MULTI_AXES = ((1,), (0, 2), (0, 1, 2), (2, 0))


class Data:
    """Seeded test data: uniform on [-1, 1) (real parts and imaginary parts independently)."""

    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)

    def complex_pairs(self, shape, dtype):
        z = self.rng.uniform(-1.0, 1.0, shape) + 1j * self.rng.uniform(-1.0, 1.0, shape)
        return to_pairs(z, dtype)

    def real(self, shape, dtype):
        return self.rng.uniform(-1.0, 1.0, shape).astype(dtype)


def to_pairs(z, dtype):
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], axis=-1), dtype=dtype)


def to_complex(p):
    p = np.asarray(p, np.float64)
    return p[..., 0] + 1j * p[..., 1]


def bound(dtype, points, x, scale):
    return (8.0 + 4.0 * math.log2(points + 1)) * EPS[dtype] * float(np.linalg.norm(np.asarray(x, np.float64))) * scale


def scale(norm, forward, points):
    if norm == "ortho":
        return 1.0 / math.sqrt(points)
    return 1.0 / points if forward == (norm == "forward") else 1.0


def worst(got, ref):
    return float(np.max(np.abs(np.asarray(got) - np.asarray(ref)))) if np.size(ref) else 0.0
