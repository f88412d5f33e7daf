"""Served construction: Python mirror of the library's construction server.

The server (rs_server_* in include/gpurecsplit_amd.h, gpurecsplit_amd/csrc/
rs_server.cpp) accepts jobs of differing key count and leaf size at any time,
copies each job's keys into device memory on arrival, and builds ready jobs with
continuous batching: `slots` build contexts (one HIP stream each) take the oldest
ready job with every ready job of the same (leaf, bucket, rotation) and build them
as one batch, admitted only while the key pool, the slots' buffers and the batch's
working-set bound fit the HBM budget.  Each job's MPHF equals its standalone build;
a failing batch is retried job by job so only the failing job reports its error.

    with Server(slots=4) as srv:
        jid = srv.submit(keys, leaf=16, bucket=2000)
        rs, info = srv.wait(jid)           # GPURecSplit, timing of the job
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import GPURecSplit, RecSplitError, _check, _flags, _keys_array, _live, lib

__all__ = ["Server", "Job", "poisson_stream", "run_stream"]

JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}
RS_ERR_PENDING = 12


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("max_batch_keys", C.c_uint64), ("time_limit_s", C.c_double)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("status", C.c_int), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("n", C.c_uint64), ("leaf", C.c_uint32), ("bucket", C.c_uint64),
                ("batch_id", C.c_uint64), ("batch_jobs", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"state": JOB_STATES.get(self.state, self.state), "status": self.status, "t_submit": self.t_submit,
                "t_start": self.t_start, "t_done": self.t_done, "n": int(self.n), "leaf": int(self.leaf),
                "bucket": int(self.bucket), "batch_id": int(self.batch_id), "batch_jobs": int(self.batch_jobs)}


class ServeStats(C.Structure):
    _fields_ = [(name, C.c_uint64) for name in (
        "jobs_submitted", "jobs_done", "jobs_failed", "keys_done", "batches", "batches_split", "committed_bytes",
        "peak_committed_bytes", "key_pool_bytes", "key_pool_used_bytes", "jobs_waiting_for_memory", "jobs_ready")
    ] + [("uptime_s", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp, _u64p = C.c_void_p, C.POINTER(C.c_uint64)
for _name, (_res, _args) in {
    "rs_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "rs_server_destroy": (None, [_vp]),
    "rs_submit_job": (C.c_int, [_vp, _u64p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, _u64p]),
    "rs_submit_job_device": (C.c_int, [_vp, C.c_void_p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, _u64p]),
    "rs_submit_batch": (C.c_int, [_vp, _u64p, _u64p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, _u64p]),
    "rs_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(JobInfo)]),
    "rs_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, C.POINTER(_vp), C.POINTER(JobInfo)]),
    "rs_server_stats": (C.c_int, [_vp, C.POINTER(ServeStats)]),
}.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class Server:
    """Construction server on one GPU (rs_server_create)."""

    def __init__(self, slots: int = 4, hbm_budget_bytes: int = 64 << 30, max_batch_keys: int = 50_000_000,
                 device: int = 0, time_limit_s: float = 0.0):
        cfg = ServerConfig(device, slots, hbm_budget_bytes, max_batch_keys, time_limit_s)
        h = C.c_void_p()
        _check(lib.rs_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.device = device
        _live["server"].add(self)

    def submit(self, keys, leaf: int, bucket: int, rotate: bool = True, remap64: bool = False) -> int:
        """Submits one job; keys: (n, 2) uint64 host array (copied at once) or a
        contiguous (n, 2) 64-bit CUDA tensor (used in place: keep it alive until wait)."""
        jid = C.c_uint64()
        if hasattr(keys, "data_ptr") and getattr(keys, "is_cuda", False):
            import torch
            torch.cuda.current_stream(keys.device).synchronize()
            _check(lib.rs_submit_job_device(self.h, C.c_void_p(keys.data_ptr()), keys.shape[0], leaf, bucket,
                                            _flags(rotate, remap64), C.byref(jid)))
        else:
            k = _keys_array(keys)
            _check(lib.rs_submit_job(self.h, k.ctypes.data_as(_u64p), k.shape[0], leaf, bucket,
                                     _flags(rotate, remap64), C.byref(jid)))
        return int(jid.value)

    def poll(self, job_id: int) -> dict:
        info = JobInfo()
        _check(lib.rs_poll(self.h, job_id, C.byref(info)))
        return info.as_dict()

    def wait(self, job_id: int, timeout_s: float = -1.0):
        """(GPURecSplit, info dict) of a finished job; raises RecSplitError for a
        failed job, and TimeoutError if it is still running after timeout_s."""
        info = JobInfo()
        out = C.c_void_p()
        rc = lib.rs_wait(self.h, job_id, timeout_s, C.byref(out), C.byref(info))
        if rc == RS_ERR_PENDING:
            raise TimeoutError(f"job {job_id} still {JOB_STATES.get(info.state)}")
        d = info.as_dict()
        if rc != 0:
            err = RecSplitError(rc, (lib.rs_last_error() or b"").decode())
            err.info = d
            raise err
        rs = GPURecSplit(bucket_size=int(info.bucket), leaf_size=int(info.leaf), _handle=out)
        return rs, d

    def stats(self) -> dict:
        s = ServeStats()
        _check(lib.rs_server_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self, _destroy=lib.rs_server_destroy):
        if getattr(self, "h", None):
            _destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


@dataclass
class Job:
    job_id: int
    keys: np.ndarray
    leaf: int
    bucket: int
    arrival: float
    start: float = -1.0
    done: float = -1.0
    batch_jobs: int = 0
    result: GPURecSplit | None = None
    error: str | None = None

    @property
    def n(self) -> int:
        return int(self.keys.shape[0])

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class StreamSummary:
    jobs: int = 0
    keys: int = 0
    wall_s: float = 0.0
    latencies: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"jobs": self.jobs, "keys": self.keys, "wall_s": self.wall_s,
                "objects_per_s": self.keys / self.wall_s if self.wall_s else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_latency_s": float(lat.mean())}


def poisson_stream(num_jobs: int, rate_per_s: float, seed: int, size_lo: int, size_hi: int,
                   leaf_weights: dict[int, float], bucket_of_leaf, size_cap_of_leaf=None):
    """Synthetic job stream: Poisson arrivals, log-uniform sizes, leaf sizes drawn
    with the given weights; each job's keys are a splitmix64 stream of its own seed.
    Returns (arrival s, n, leaf, bucket, key seed) per job."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=num_jobs))
    sizes = np.exp(rng.uniform(np.log(size_lo), np.log(size_hi), size=num_jobs)).astype(np.int64)
    leaves = list(leaf_weights)
    ls = rng.choice(leaves, size=num_jobs, p=np.array([leaf_weights[l] for l in leaves]) / sum(leaf_weights.values()))
    jobs = []
    for i in range(num_jobs):
        n = int(sizes[i])
        if size_cap_of_leaf:
            n = min(n, size_cap_of_leaf(int(ls[i])))
        jobs.append((float(t[i]), n, int(ls[i]), bucket_of_leaf(int(ls[i])), int(rng.integers(1, 2**62))))
    return jobs


def run_stream(stream, slots: int = 4, hbm_budget_bytes: int = 64 << 30, time_limit_s: float = 0.0,
               on_done=None, keys_of=None) -> tuple[list[Job], dict]:
    """Replays a stream in real time against a Server: each job is submitted at
    its arrival time (keys generated beforehand), results are collected as they
    finish.  Returns (jobs, summary with p50/p99 latency = done - arrival and
    objects/s over the stream's wall time).  keys_of(job tuple) may supply keys."""
    from .keys import splitmix64_keys

    prepared = [(s[0], keys_of(s) if keys_of else splitmix64_keys(s[4], s[1]), s[2], s[3]) for s in stream]
    summ = StreamSummary()
    jobs: list[Job] = []
    with Server(slots=slots, hbm_budget_bytes=hbm_budget_bytes, time_limit_s=time_limit_s) as srv:
        t0 = time.perf_counter()
        pending = []
        for i, (t_arr, keys, leaf, bucket) in enumerate(prepared):
            delay = t_arr - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            jid = srv.submit(keys, leaf, bucket)
            j = Job(jid, keys, leaf, bucket, arrival=time.perf_counter() - t0)
            jobs.append(j)
            pending.append(j)
        for j in pending:
            try:
                rs, info = srv.wait(j.job_id)
                j.result = rs
            except RecSplitError as e:
                info = e.info
                j.error = str(e)
            # the server clock starts at Server creation, slightly before t0
            skew = info["t_submit"] - j.arrival
            j.start, j.done = info["t_start"] - skew, info["t_done"] - skew
            j.batch_jobs = info["batch_jobs"]
            summ.jobs += 1
            summ.keys += j.n
            summ.latencies.append(j.latency)
            if on_done is not None:
                on_done(j)
        summ.wall_s = max(j.done for j in jobs) if jobs else 0.0
        out = summ.summary()
        out["server"] = srv.stats()
    out["batches"] = out["server"]["batches"]
    return jobs, out
