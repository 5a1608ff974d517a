"""GPU: the accumulation primitives (include/aether/device/accum.hpp) through their bulk entries.

Keyed sums follow the reference's own cases (tests/test_Run_common.h, tests/test_Run.cu) in wave64 form:
every case is a keyed sum of terms in [-1, 1) whose error against a long-double reference, over
eps * sum|t| per slot, is within 8 plain and 1 compensated (8 when runs are shorter than a wave), and
whose commit count is the one the wave's run structure implies. Planes: concurrent producers on four
streams deliver the exact sum to within one rounding; the serialized policy is the plain sequential
sum. Append counters: every accepted claim gets a distinct slot below capacity, the set of appended
indices is the set asked for, overflow is flagged and clamped (tests/test_AppendCounter.cu).
"""
import math

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import accum as A

pytestmark = pytest.mark.gpu
WAVE = 64


def unit_of(l):
    """A uniform value in [0, 1) for lane l (tests/test_Run_common.h:44-51)."""
    x = ((l.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
    x ^= x >> np.uint64(31)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(29)
    return (x >> np.uint64(11)).astype(np.float64) * 2.0**-53


def by_quotient(n, run, own=0):
    l = np.arange(n, dtype=np.int64)
    return (l + own) // run


def mixed_runs(n):
    keys = np.empty(n, np.int64)
    k, left = 0, 1
    u = unit_of(np.arange(n))
    for l in range(n):
        left -= 1
        if left == 0:
            k += 1
            left = 1 + int(u[l] * 1000.0) % 40
        keys[l] = k
    return keys


def expected_commits(keys, form, width, own):
    """(run, lane) commits of the wave64 forms for n a multiple of 64."""
    n = len(keys)
    w = keys.reshape(-1, WAVE)
    uniform = np.all(w == w[:, :1], axis=1)
    heads = 1 + np.count_nonzero(w[:, 1:] != w[:, :-1], axis=1)
    waves = n // WAVE
    if form == A.ALIGNED:
        return (waves * (WAVE // width), 0) if own % WAVE == 0 else (0, n)
    if form == A.TESTED:
        return int(uniform.sum()), int((~uniform).sum()) * WAVE
    return int(heads.sum()), 0


def run_case(keys, form, width=64, own=0, short_runs=False):
    n = len(keys)
    slots = int(keys.max()) + 1
    t64 = 2.0 * unit_of(np.arange(n)) - 1.0
    want = expected_commits(keys, form, width, own)
    for dt, np_dt in ((torch.float32, np.float32), (torch.float64, np.float64)):
        t = t64.astype(np_dt)
        ref = np.zeros(slots, np.longdouble)
        mass = np.zeros(slots, np.longdouble)
        np.add.at(ref, keys, t.astype(np.longdouble))
        np.add.at(mass, keys, np.abs(t).astype(np.longdouble))
        eps = float(np.finfo(np_dt).eps)
        for comp in (False, True):
            tg, cp, commits = A.keyed_sum(torch.from_numpy(keys).cuda(), torch.from_numpy(t).cuda(), slots, form,
                                          width, comp, own, count_commits=True)
            torch.cuda.synchronize()
            got = tg.cpu().numpy().astype(np.longdouble)
            if comp:
                got = got + cp.cpu().numpy().astype(np.longdouble)
            m = mass > 0
            worst = float(np.max(np.abs(got[m] - ref[m]) / (eps * mass[m])))
            bound = 1.0 if comp and not short_runs else 8.0
            what = f"form {form} width {width} {'comp' if comp else 'plain'} {dt}"
            assert worst <= bound, what
            assert tuple(int(v) for v in commits.cpu()) == want, what


def test_aligned_run_commits_once_per_wave(dev):
    run_case(by_quotient(64 * 8192, 8192), A.ALIGNED, 64)


def test_sub_wave_aligned_runs_commit_once_per_group(dev):
    run_case(by_quotient(1024 * 48, 48), A.ALIGNED, 16)
    run_case(by_quotient(1024 * 64, 32), A.ALIGNED, 32)


def test_misaligned_wave_commits_per_lane(dev):
    run_case(by_quotient(64 * 1024, 32, own=1000), A.ALIGNED, 32, own=1000)


def test_tested_run_commits_once_per_uniform_wave(dev):
    run_case(by_quotient(64 * 8192, 64), A.TESTED)
    run_case(by_quotient(64 * 1024, 8), A.TESTED, short_runs=True)


def test_sorted_run_commits_once_per_run_in_wave(dev):
    run_case(by_quotient(64 * 8192, 8192), A.SORTED)
    run_case(mixed_runs(64 * 8192), A.SORTED, short_runs=True)


def test_unsorted_keys_stay_correct(dev):
    keys = (np.arange(64 * 1024, dtype=np.int64) * 7919) % 1024
    run_case(keys, A.SORTED, short_runs=True)


def test_out_of_range_keys_take_no_part(dev):
    keys = torch.tensor([0, 1, 5, -1, 2], dtype=torch.int64, device="cuda")
    terms = torch.ones(5, dtype=torch.float64, device="cuda")
    tg, _, _ = A.keyed_sum(keys, terms, 3, "sorted")
    assert tg.cpu().tolist() == [1.0, 1.0, 1.0]


def test_plane_concurrent_producers_deliver_the_exact_sum(dev):
    n, vd, k = 1 << 16, 3, 8
    plane = A.Plane(n, vd)
    terms = []
    for s in range(k):
        x = ae.Array(n, vd, torch.float64)
        ae.array_fill_normal(x, 1000 + s, 0.0, 10.0 ** (s % 4))  # magnitudes that cancel badly
        terms.append(x)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(4)]
    for s, x in enumerate(terms):
        plane.add(x, stream=streams[s % 4])
    for st in streams:
        st.synchronize()
    got = plane.delivered().data[:, :n].cpu().numpy()
    T = np.stack([x.data[:, :n].cpu().numpy() for x in terms]).reshape(k, -1)
    exact = np.array([math.fsum(T[:, j]) for j in range(T.shape[1])]).reshape(vd, n)  # correctly rounded
    ulp = np.spacing(np.abs(exact))
    assert np.max(np.abs(got - exact) / ulp) <= 1.0
    # the same producers in the reverse order on other streams: within the same rounding
    plane2 = A.Plane(n, vd)
    for s, x in enumerate(reversed(terms)):
        plane2.add(x, stream=streams[(s + 1) % 4])
    torch.cuda.synchronize()
    got2 = plane2.delivered().data[:, :n].cpu().numpy()
    assert np.max(np.abs(got2 - got) / ulp) <= 1.0


def _body_terms(cancel):
    """The value families of tests/test_AccumOrderPermutation.cu:90-115 (8 bodies x 3 components x 128
    samples): magnitudes spread over eleven decades with hashed signs, or bodies that cancel pairwise to
    within 0 .. 1e-3 relative; each value carries the same hashed jitter."""
    kb, ks = 8, 128
    i = np.arange(ks, dtype=np.uint64)
    v = np.zeros((kb, 3, ks))
    rel = [1.0, 1.1e-3, 3.1e-6, 6.0e-7, 5.0e-8, 1.0e-10, 4.0e-12, 8.0e-13]
    dl = np.array([0.0, 2.22e-16, 9.1e-13, 9.3e-10, 9.5e-7, 9.8e-4])[(i % 6).astype(np.int64)]
    for k in range(kb):
        for d in range(3):
            h = (i * np.uint64(2654435761) + np.uint64(d * 40503) + np.uint64(k * 2246822519)) & np.uint64(0xFFFFFFFF)
            jit = 1.0 + 0.4375 * ((h >> np.uint64(8)) & np.uint64(0xFFFF)).astype(np.float64) / 65536.0
            if cancel:
                a = 1.0e-2 * jit
                v[k, d] = a if k % 2 == 0 else -a * (1.0 - dl)
            else:
                sgn = np.where((h >> np.uint64(3)) & np.uint64(1), 1.0, -1.0)
                v[k, d] = 8.0e-3 * rel[k] * jit * (1.0 if k == 0 else sgn)
    return v


@pytest.mark.parametrize("cancel", [False, True], ids=["spread", "cancelling"])
def test_plane_bitwise_under_permuted_launch_orders(dev, cancel):
    """The reference's admissibility bar for the compensated plane (tests/test_AccumOrderPermutation.cu:182):
    the same 8 bodies added one launch each in five orders (identity, reverse, three shuffles) deliver planes
    that are bit-identical, and equal to the correctly rounded sum of each sample's 8 terms."""
    v = _body_terms(cancel)
    bodies = []
    for k in range(v.shape[0]):
        x = ae.Array(v.shape[2], 3, torch.float64)
        x.data[:, : v.shape[2]] = torch.from_numpy(v[k]).to("cuda")
        bodies.append(x)
    rng = np.random.default_rng(0xACC0)
    orders = [list(range(8)), list(range(7, -1, -1))] + [list(rng.permutation(8)) for _ in range(3)]
    planes = []
    for order in orders:
        plane = A.Plane(v.shape[2], 3)
        for k in order:
            plane.add(bodies[k])
        planes.append(plane.delivered().data[:, : v.shape[2]].cpu().numpy().copy())
    for p in planes[1:]:
        assert np.array_equal(p.view(np.uint64), planes[0].view(np.uint64))
    exact = np.array([[math.fsum(v[:, d, j]) for j in range(v.shape[2])] for d in range(3)])
    assert np.array_equal(planes[0], exact)


def test_serialized_plane_is_the_sequential_sum(dev):
    n, vd = 100_003, 2
    plane = A.Plane(n, vd, policy="serialized")
    terms = []
    for s in range(3):
        x = ae.Array(n, vd, torch.float64)
        ae.array_fill_uniform(x, 50 + s, -1.0, 1.0)
        terms.append(x)
        plane.add(x)
    torch.cuda.synchronize()
    acc = np.zeros((vd, n))
    for x in terms:
        acc = acc + x.data[:, :n].cpu().numpy()
    assert np.array_equal(plane.delivered().data[:, :n].cpu().numpy(), acc)
    plane.zero()
    torch.cuda.synchronize()
    assert torch.count_nonzero(plane.value.data[:, :n]).item() == 0


@pytest.mark.parametrize("claim", ["tile", "wave", "lane"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_append_where_hands_out_every_slot_once(dev, claim, dt):
    n = 1_000_003
    x = torch.rand(n, dtype=torch.float64, generator=torch.Generator().manual_seed(3)).to(dt).cuda()
    want = np.nonzero(x.cpu().numpy().astype(np.float64) > 0.9)[0]
    ctr = A.AppendCounter(200_000)
    out = torch.full((ctr.capacity,), -1, dtype=torch.int32, device="cuda")
    A.append_where(x, 0.9, ctr, out, claim)
    torch.cuda.synchronize()
    assert ctr.raw_count() == len(want) and not ctr.overflowed()
    k = ctr.commit()
    got = out[:k].cpu().numpy()
    assert k == len(want) and np.array_equal(np.sort(got), want)
    assert ctr.raw_count() == 0  # commit reset the counter


@pytest.mark.parametrize("claim", ["tile", "wave", "lane"])
def test_append_overflow_is_flagged_and_clamped(dev, claim):
    n = 100_000
    x = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device="cuda")
    ctr = A.AppendCounter(1000)
    out = torch.full((1000,), -1, dtype=torch.int32, device="cuda")
    A.append_where(x, 0.5, ctr, out, claim)
    torch.cuda.synchronize()
    total = int((x > 0.5).sum().item())
    assert ctr.raw_count() == total and ctr.overflowed()
    k = ctr.commit()
    got = out[:k].cpu().numpy()
    assert k == 1000 and len(np.unique(got)) == 1000 and np.all(x.cpu().numpy()[got] > 0.5)
    with pytest.raises(ae.AetherError, match="DeviceFlag"):
        ctr.check_overflow()
    assert not ctr.overflowed()


def test_append_tile_keeps_order_within_a_tile(dev):
    # one tile's claimants take consecutive slots in element order
    x = torch.zeros(16384, dtype=torch.float64, device="cuda")
    x[::3] = 1.0
    ctr = A.AppendCounter(16384)
    out = torch.full((16384,), -1, dtype=torch.int32, device="cuda")
    A.append_where(x, 0.5, ctr, out, "tile")
    k = ctr.commit()
    assert np.array_equal(out[:k].cpu().numpy(), np.arange(0, 16384, 3))
