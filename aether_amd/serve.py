"""Python binding of the job server (aeth_server_* in include/aether_hip.h).

Also generates the served mixed job stream: Poisson arrivals; 50% fused
evaluations (samples log-uniform 2^10..2^24, FP32/FP64 50/50), 25% Philox
sampling jobs (2^12..2^26 normals), 25% FFT batches (1..64 signals of
lengths 2^6..2^13), all from fixed seeds.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib

JOB_EVAL, JOB_SAMPLE, JOB_FFT = 0, 1, 2
SERVER_GRAPH_BATCHES = 1
SERVER_TUNE = 2  # with SERVER_GRAPH_BATCHES: tune the graph form's persistent grids online
F64, F32 = 0, 1


class Job(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("dtype", C.c_uint32), ("n", C.c_uint64), ("seed", C.c_uint64),
                ("arrival", C.c_double), ("fft_log2n", C.c_uint8 * 64)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("kind", C.c_uint32), ("status", C.c_int32), ("value", C.c_double * 8),
                ("submit_time", C.c_double), ("done_time", C.c_double)]


class Config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("streams", C.c_uint32), ("max_batch_jobs", C.c_uint32),
                ("fp32_compensated", C.c_uint32), ("flags", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("pool_bytes_in_use", C.c_uint64), ("pool_bytes_budget", C.c_uint64)]


def _bind():
    L = lib()
    sig = {
        "aeth_server_create": [C.POINTER(Config), C.POINTER(C.c_void_p)],
        "aeth_server_submit": [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_uint64)],
        "aeth_server_poll": [C.c_void_p, C.POINTER(JobResult), C.c_uint32, C.POINTER(C.c_uint32)],
        "aeth_server_wait_idle": [C.c_void_p],
        "aeth_server_get_stats": [C.c_void_p, C.POINTER(Stats)],
        "aeth_server_tuning": [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
        "aeth_server_destroy": [C.c_void_p],
        "aeth_server_replay": [C.c_void_p, C.POINTER(Job), C.c_uint64, C.POINTER(JobResult)],
        "aeth_run_job_plain": [C.POINTER(Job), C.POINTER(JobResult), C.c_void_p],
        "aeth_run_job_plain_ws": [C.POINTER(Job), C.POINTER(JobResult), C.c_void_p, C.c_uint64, C.c_void_p],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = args
    return L


def make_job(kind, n, seed, dtype=F64, fft_log2n=(), arrival=0.0) -> Job:
    j = Job()
    j.kind, j.dtype, j.n, j.seed, j.arrival = kind, dtype, n, seed, arrival
    for k, v in enumerate(fft_log2n):
        j.fft_log2n[k] = v
    if kind == JOB_FFT:
        j.n = len(fft_log2n)
    return j


def make_mix(rate: float, duration: float, seed: int = 5):
    """The served mixed stream as a ctypes array of Job (arrival-ordered)."""
    rng = np.random.default_rng(seed)
    count = int(rng.poisson(rate * duration))
    arrivals = np.sort(rng.uniform(0.0, duration, count))
    kinds = rng.choice([JOB_EVAL, JOB_SAMPLE, JOB_FFT], size=count, p=[0.5, 0.25, 0.25])
    jobs = (Job * count)()
    for i in range(count):
        k = int(kinds[i])
        j = jobs[i]
        j.kind, j.arrival, j.seed = k, float(arrivals[i]), int(rng.integers(1, 2**40)) * 8
        if k == JOB_EVAL:
            j.dtype = F32 if rng.random() < 0.5 else F64
            j.n = int(2.0 ** rng.uniform(10, 24))
        elif k == JOB_SAMPLE:
            j.dtype = F64
            j.n = int(2.0 ** rng.uniform(12, 26))
        else:
            j.dtype = F64
            ns = int(rng.integers(1, 65))
            j.n = ns
            for q in range(ns):
                j.fft_log2n[q] = int(rng.integers(6, 14))
    return jobs


def job_bytes(j: Job, fp32_compensated: bool = False) -> float:
    """Algorithmic bytes of a job's hot-path work in the precision the server executes it in (a server
    in the FP32-compensated form runs every job in FP32 storage): eval 3 loads + 1 store per
    (component, sample); sampling: bytes written; FFT: read and write once per point."""
    es = 8 if (j.dtype == F64 and not fp32_compensated) else 4
    if j.kind == JOB_EVAL:
        return 3 * j.n * 4 * es
    if j.kind == JOB_SAMPLE:
        return j.n * es
    pts = sum(1 << j.fft_log2n[k] for k in range(j.n))
    return 2 * pts * 2 * es


class Server:
    def __init__(self, budget_bytes: int = 64 << 30, streams: int = 4, max_batch_jobs: int = 256,
                 fp32_compensated: bool = False, graph_batches: bool = False, tune: bool = False):
        self.L = _bind()
        flags = (SERVER_GRAPH_BATCHES if graph_batches else 0) | (SERVER_TUNE if tune else 0)
        cfg = Config(budget_bytes, streams, max_batch_jobs, int(fp32_compensated), flags)
        h = C.c_void_p()
        call("aeth_server_create", C.byref(cfg), C.byref(h))
        self.h = h

    def submit(self, job: Job) -> int:
        jid = C.c_uint64()
        call("aeth_server_submit", self.h, C.byref(job), C.byref(jid))
        return jid.value

    def wait_idle(self):
        call("aeth_server_wait_idle", self.h)

    def poll(self, max_results: int = 4096):
        buf = (JobResult * max_results)()
        n = C.c_uint32()
        call("aeth_server_poll", self.h, buf, max_results, C.byref(n))
        return list(buf[: n.value])

    def replay(self, jobs):
        res = (JobResult * len(jobs))()
        call("aeth_server_replay", self.h, jobs, len(jobs), res)
        return res

    def stats(self) -> Stats:
        s = Stats()
        call("aeth_server_get_stats", self.h, C.byref(s))
        return s

    def tuning(self) -> dict:
        """The graph form's persistent grid in blocks per CU, and whether the online tuner has settled."""
        b, done = C.c_uint32(), C.c_uint32()
        call("aeth_server_tuning", self.h, C.byref(b), C.byref(done))
        return {"blocks_per_cu": b.value, "settled": bool(done.value)}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            call("aeth_server_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_plain(job: Job, stream: int = 0) -> JobResult:
    _bind()
    r = JobResult()
    call("aeth_run_job_plain", C.byref(job), C.byref(r), stream)
    return r


def run_plain_ws(job: Job, workspace: int, workspace_bytes: int, stream: int = 0) -> JobResult:
    """The plain per-call form with the job's arrays in a caller-owned device workspace."""
    _bind()
    r = JobResult()
    call("aeth_run_job_plain_ws", C.byref(job), C.byref(r), workspace, workspace_bytes, stream)
    return r
