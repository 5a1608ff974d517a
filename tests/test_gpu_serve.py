"""GPU: the job server. Each result is bit for bit the same alone (plain
per-call path) and inside a served mixed batch; results match the oracle;
admission respects the HBM budget.
"""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import serve as S

pytestmark = pytest.mark.gpu


def jobs_small_mix(count=60, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = i % 3
        if k == 0:
            out.append(S.make_job(S.JOB_EVAL, int(2 ** rng.uniform(5, 18)), 1000 + 8 * i,
                                  dtype=S.F32 if i % 2 else S.F64))
        elif k == 1:
            out.append(S.make_job(S.JOB_SAMPLE, int(2 ** rng.uniform(5, 19)), 5000 + 8 * i,
                                  dtype=S.F32 if i % 4 == 1 else S.F64))
        else:
            lens = [int(v) for v in rng.integers(0, 14, int(rng.integers(1, 9)))]
            out.append(S.make_job(S.JOB_FFT, len(lens), 9000 + 8 * i, dtype=S.F64 if i % 2 else S.F32,
                                  fft_log2n=lens))
    return out


def test_alone_equals_in_the_mix(dev):
    jobs = jobs_small_mix()
    alone = [S.run_plain(j) for j in jobs]
    with S.Server(budget_bytes=4 << 30, streams=4, max_batch_jobs=64) as srv:
        ids = [srv.submit(j) for j in jobs]
        srv.wait_idle()
        got = {r.id: r for r in srv.poll()}
        st = srv.stats()
    assert st.completed == len(jobs) and st.failed == 0
    assert st.batches < len(jobs)  # jobs were coalesced
    for j, a, i in zip(jobs, alone, ids):
        r = got[i]
        assert r.status == 0
        assert list(r.value) == list(a.value), (j.kind, j.dtype, j.n, list(r.value), list(a.value))
        used = {S.JOB_EVAL: 6, S.JOB_SAMPLE: 4, S.JOB_FFT: 2}[j.kind]
        assert all(v == 0.0 for v in list(r.value)[used:])


@pytest.mark.parametrize("fp32", [False, True], ids=["requested", "fp32_compensated"])
def test_graph_batches_equal_direct_launches(dev, fp32):
    """A server whose batches are replays of one captured graph per stream slot (persistent kernels reading
    the batch's counts on the device) gives every job the bits of the direct launches."""
    jobs = jobs_small_mix(count=48, seed=5)
    results = {}
    for graph in (False, True):
        with S.Server(budget_bytes=4 << 30, streams=2, max_batch_jobs=16, fp32_compensated=fp32,
                      graph_batches=graph) as srv:
            ids = [srv.submit(j) for j in jobs]
            srv.wait_idle()
            got = {r.id: r for r in srv.poll()}
            st = srv.stats()
        assert st.completed == len(jobs) and st.failed == 0 and st.batches >= len(jobs) // 16
        results[graph] = [got[i] for i in ids]
    for j, a, b in zip(jobs, results[False], results[True]):
        assert a.status == 0 and b.status == 0
        assert list(a.value) == list(b.value), (j.kind, j.dtype, j.n)


def test_tuned_graph_batches_settle_and_equal_direct_launches(dev):
    """The graph form's online tuner (AETH_SERVER_TUNE) times batches over four persistent grids, settles on
    one of them after 8 timed batches each, and every job keeps the bits of the direct launches."""
    jobs = jobs_small_mix(count=120, seed=13)
    results = {}
    for tune in (False, True):
        with S.Server(budget_bytes=4 << 30, streams=2, max_batch_jobs=3, graph_batches=tune, tune=tune) as srv:
            ids = [srv.submit(j) for j in jobs]
            srv.wait_idle()
            got = {r.id: r for r in srv.poll()}
            st, tn = srv.stats(), srv.tuning()
        assert st.completed == len(jobs) and st.failed == 0
        results[tune] = [got[i] for i in ids]
        if tune:
            assert st.batches >= 32 and tn["settled"] and tn["blocks_per_cu"] in (2, 4, 8, 16), (st.batches, tn)
        else:
            assert tn == {"blocks_per_cu": 8, "settled": False}
    for j, a, b in zip(jobs, results[False], results[True]):
        assert a.status == 0 and b.status == 0
        assert list(a.value) == list(b.value), (j.kind, j.dtype, j.n)


def test_eval_job_matches_oracle(dev, orc):
    n, seed = 10_000, 77
    r = S.run_plain(S.make_job(S.JOB_EVAL, n, seed, S.F64))
    pitch = -(-n // 32) * 32
    A, B, C = (orc.fill_array_uniform(seed + k, 3, n, pitch) for k in range(3))
    d = orc.eval_cfg0(A, B, C, 3, n, pitch, mathmode=1).reshape(3, pitch)[:, :n]
    for k in range(3):
        h, c = orc.device_sum(np.ascontiguousarray(d[k]))
        assert (r.value[2 * k], r.value[2 * k + 1]) == (h, c)


def test_sample_job_matches_oracle(dev, orc):
    n, seed = 50_001, 99
    r = S.run_plain(S.make_job(S.JOB_SAMPLE, n, seed, S.F64))
    z = orc.fill_normal_f64(seed, n)
    assert (r.value[0], r.value[1]) == orc.device_sum(z)
    assert (r.value[2], r.value[3]) == orc.device_sum(z, square=True)


def test_fft_job_matches_oracle_within_bound(dev, orc):
    lens = [6, 10, 3]
    seed = 1234
    r = S.run_plain(S.make_job(S.JOB_FFT, 3, seed, S.F64, fft_log2n=lens))
    total = sum(1 << l for l in lens)
    x = orc.fill_normal_f64(seed, 2 * total).view(np.complex128)
    pos, acc = 0, 0.0
    for l in lens:
        acc += orc.dft(x[pos:pos + (1 << l)]).view(np.float64).sum()
        pos += 1 << l
    got = r.value[0] + r.value[1]
    assert abs(got - acc) <= 1e-9 * np.abs(x).sum()


def test_budget_admission_and_oversize(dev):
    jobs = [S.make_job(S.JOB_EVAL, 1 << 16, 8 * i + 1, S.F64) for i in range(20)]
    big = S.make_job(S.JOB_EVAL, 1 << 22, 7, S.F64)  # needs ~400 MB > budget
    with S.Server(budget_bytes=32 << 20, streams=2, max_batch_jobs=8) as srv:
        ids = [srv.submit(j) for j in jobs]
        bid = srv.submit(big)
        srv.wait_idle()
        res = {r.id: r for r in srv.poll()}
    assert all(res[i].status == 0 for i in ids)
    assert res[bid].status == 2


def test_fp32_compensated_mode_error(dev):
    """A server in the FP32-compensated form evaluates in FP32 and accumulates in float-float. Its sums
    are held against an FP64 compensated reduction of the same float values of d (recomputed by the
    single-call FP32 kernel, which the served batch matches tile for tile) to 2^-44 sum|d| -- a bound that
    a plain FP32 sum of those values does not meet."""
    n = 1 << 15
    jobs = [S.make_job(S.JOB_EVAL, n, 8 * i + 3, S.F64) for i in range(6)]
    with S.Server(budget_bytes=1 << 30, fp32_compensated=True) as srv:
        ids = [srv.submit(j) for j in jobs]
        srv.wait_idle()
        res = {r.id: r for r in srv.poll()}
    plain_errors, bounds = [], []
    for j, i in zip(jobs, ids):
        a, b, c, d = (ae.Array(n, 3, torch.float32) for _ in range(4))
        for k, x in enumerate((a, b, c)):
            ae.array_fill_uniform(x, j.seed + k, -1.0, 1.0)
        plain = ae.eval_cfg0(a, b, c, d, ae.SUM_PLAIN).double().sum(1).cpu()
        d64 = ae.Array(n, 3, torch.float64)
        d64.data[:, :n].copy_(d.data[:, :n])
        ref = ae.reduce(d64, ae.SUM_NEUMAIER).sum(1).cpu()
        abs_sum = d64.data[:, :n].abs().sum(1).cpu()
        for k in range(3):
            served = res[i].value[2 * k] + res[i].value[2 * k + 1]
            bound = 2.0**-44 * abs_sum[k].item()
            assert abs(served - ref[k].item()) <= bound, (k, served, ref[k].item())
            plain_errors.append(abs(plain[k].item() - ref[k].item()))
            bounds.append(bound)
    assert max(e / b for e, b in zip(plain_errors, bounds)) > 1.0, "plain FP32 would pass: the bound shows nothing"


def test_replay_open_loop(dev):
    jobs = S.make_mix(rate=2000.0, duration=0.2, seed=11)
    with S.Server(budget_bytes=16 << 30) as srv:
        res = srv.replay(jobs)
    assert all(r.status == 0 for r in res)
    lat = np.array([r.done_time - r.submit_time for r in res])
    assert np.all(lat > 0)
