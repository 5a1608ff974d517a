"""GPU: the hot path captured in a hipGraph and replayed.

A stream capture records the Philox array fills, the fused evaluation (the specialised kernel and the
runtime-compiled tape), a Philox bit stream and a batched uneven FFT; every replay must reproduce the eager
results bit for bit. Between replays the outputs are cleared by eager work on another stream and the
program synchronises, so each replay is checked on its own. Plans and runtime-compiled kernels are built
before the capture begins (building either is not allowed inside one)."""
import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import expr as E

pytestmark = pytest.mark.gpu


def outputs(d, d2, bits, y):
    torch.cuda.synchronize()
    return [d.data[:, : d.n].clone(), d2.data[:, : d2.n].clone(), bits.clone(), y.clone()]


def test_hot_path_in_a_graph_replays_bit_for_bit(dev):
    n = 100_003
    s = torch.cuda.Stream()
    a, b, c, d, d2 = (ae.Array(n, 3, torch.float64) for _ in range(5))
    bits = torch.empty(1 << 20, dtype=torch.int32, device="cuda")
    lengths = np.array([64, 8192, 512, 4096, 128, 2048], dtype=np.uint64)
    total = int(lengths.sum())
    g0 = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(total, dtype=torch.complex128, generator=g0).to("cuda")
    y = torch.empty_like(x)
    plan = ae.FftBatch(lengths, prec=0)
    expr = a * b + E.sin(c) - E.atan(a * 2.0)

    def work(stream):
        for arr, seed in ((a, 42), (b, 43), (c, 44)):
            ae.array_fill_uniform(arr, seed, -1.0, 1.0, stream=stream)
        ae.eval_cfg0(a, b, c, d, ae.SUM_NONE, stream=stream)
        E.assign(d2, expr, stream=stream)
        ae.rng_fill(ae.RNG_BITS32, 7, bits, stream=stream)
        plan(torch.view_as_real(x), torch.view_as_real(y), True, 0, stream=stream)

    # eager reference (also compiles the tape's kernel before the capture)
    with torch.cuda.stream(s):
        work(s)
    s.synchronize()
    want = outputs(d, d2, bits, y)

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        work(s)
    for _ in range(3):
        for t in (d.data, d2.data, bits, y):
            t.zero_()
        torch.cuda.synchronize()
        g.replay()
        got = outputs(d, d2, bits, y)
        for w, h in zip(want, got):
            assert torch.equal(w, h)
    plan.close()


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_fused_sums_in_a_graph_replay(dev, dt):
    """The fused evaluation with its compensated sums (the partials live in stream-ordered scratch, which a
    capture records as allocation nodes) replays to the eager sums bit for bit."""
    n = 300_001
    s = torch.cuda.Stream()
    a, b, c, d = (ae.Array(n, 3, dt) for _ in range(4))
    mode = ae.SUM_NEUMAIER if dt == torch.float64 else ae.SUM_FLOATFLOAT
    sums = torch.empty((3, 2), dtype=dt, device="cuda")
    sums2 = torch.empty((3, 2), dtype=dt, device="cuda")

    def work(stream):
        for arr, seed in ((a, 5), (b, 6), (c, 7)):
            ae.array_fill_uniform(arr, seed, -1.0, 1.0, stream=stream)
        ae.eval_cfg0(a, b, c, d, mode, sums, stream=stream)
        E.assign(d, a + b * c + E.sin(a) * E.exp(-1.0 * (b * b)), mode, sums2, stream=stream)

    with torch.cuda.stream(s):
        work(s)
    s.synchronize()
    want, want2 = sums.clone(), sums2.clone()
    assert torch.equal(want, want2)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        work(s)
    for _ in range(3):
        sums.zero_()
        sums2.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(sums, want) and torch.equal(sums2, want)


def test_served_batches_replay_one_graph(dev):
    """Served batches as replays of a graph captured once per stream slot: three waves of jobs, with eager
    GPU work and synchronisations between them, each job's result equal to the plain per-call path."""
    from aether_amd import serve as S

    rng = np.random.default_rng(17)
    waves = []
    for w in range(3):
        wave = []
        for i in range(12):
            seed = 100 * w + 8 * i + 1
            if i % 3 == 0:
                wave.append(S.make_job(S.JOB_EVAL, int(2 ** rng.uniform(8, 16)), seed, S.F64))
            elif i % 3 == 1:
                wave.append(S.make_job(S.JOB_SAMPLE, int(2 ** rng.uniform(8, 17)), seed, S.F32))
            else:
                lens = [int(v) for v in rng.integers(2, 14, 3)]
                wave.append(S.make_job(S.JOB_FFT, len(lens), seed, S.F64, fft_log2n=lens))
        waves.append(wave)
    busy = torch.ones(1 << 20, device="cuda")
    with S.Server(budget_bytes=2 << 30, streams=1, max_batch_jobs=4, graph_batches=True) as srv:
        for wave in waves:
            ids = [srv.submit(j) for j in wave]
            srv.wait_idle()
            got = {r.id: r for r in srv.poll()}
            busy.mul_(1.0001).add_(0.5)  # the program's own eager work between replays
            torch.cuda.synchronize()
            for j, i in zip(wave, ids):
                assert got[i].status == 0
                assert list(got[i].value) == list(S.run_plain(j).value), (j.kind, j.n)
        assert srv.stats().batches >= 9  # 36 jobs, at most 4 a batch: the one graph replayed again and again
