"""GPU: the allocation-free form of the fused evaluation (aeth_eval_cfg0_noalloc) against
aeth_eval_cfg0, through the C ABI: the same output and sums bit for bit on two calls in a row, the
second with a smaller input (the shared scratch keeps the first call's contents); sizes whose sum
partials exceed the fixed scratch are refused with an error."""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import _lib

pytestmark = pytest.mark.gpu


def _noalloc():
    return _lib.lib().aeth_eval_cfg0_noalloc


def _run(fn, dt, arrays, mode, sums, st):
    a, b, c, d = arrays
    code = ae.F64 if dt == torch.float64 else ae.F32
    return fn(code, a.data.data_ptr(), b.data.data_ptr(), c.data.data_ptr(), d.data.data_ptr(), 3, a.n, a.capacity,
              mode, sums.data_ptr() if sums is not None else None, st)


@pytest.mark.parametrize("dt,mode", [(torch.float64, ae.SUM_NEUMAIER), (torch.float64, ae.SUM_PLAIN),
                                     (torch.float32, ae.SUM_FLOATFLOAT), (torch.float32, ae.SUM_NEUMAIER),
                                     (torch.float64, ae.SUM_NONE)])
def test_noalloc_equals_allocating_entry(dev, dt, mode):
    fn = _noalloc()
    ref_fn = _lib.lib().aeth_eval_cfg0
    st = torch.cuda.current_stream().cuda_stream
    for n, seed in ((5_000_003, 42), (100_003, 7)):  # second call smaller: stale scratch must not leak in
        arrays = [ae.Array(n, 3, dt) for _ in range(4)]
        for arr, s in zip(arrays[:3], (seed, seed + 1, seed + 2)):
            ae.array_fill_uniform(arr, s)
        d_ref = ae.Array(n, 3, dt)
        s_ref = torch.zeros(6, dtype=dt, device="cuda") if mode != ae.SUM_NONE else None
        s_new = torch.full((6,), float("nan"), dtype=dt, device="cuda") if mode != ae.SUM_NONE else None
        assert _run(ref_fn, dt, arrays[:3] + [d_ref], mode, s_ref, st) == 0
        assert _run(fn, dt, arrays, mode, s_new, st) == 0, _lib.lib().aeth_last_error()
        torch.cuda.synchronize()
        # samples past n are padding no entry writes
        assert torch.equal(arrays[3].data[:, :n], d_ref.data[:, :n])
        if mode != ae.SUM_NONE:
            assert np.array_equal(s_new.cpu().numpy().view(np.uint64 if dt == torch.float64 else np.uint32),
                                  s_ref.cpu().numpy().view(np.uint64 if dt == torch.float64 else np.uint32))


def test_noalloc_refuses_sizes_beyond_its_scratch(dev):
    fn = _noalloc()
    # 3 x 2^28 float64 samples need 393,216 tile partials (6 MiB) > the 4 MiB scratch; the check happens
    # before any launch, so the planes are never touched (null pointers are never dereferenced)
    n = 1 << 28
    sums = torch.zeros(6, dtype=torch.float64, device="cuda")
    rc = fn(ae.F64, None, None, None, None, 3, n, n, ae.SUM_NEUMAIER, sums.data_ptr(), None)
    assert rc == 1
    assert b"fixed scratch" in _lib.lib().aeth_last_error()
