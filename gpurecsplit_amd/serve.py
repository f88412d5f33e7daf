"""Served construction: a request loop with continuous batching on one GPU.

Jobs (key set, leaf size, bucket size) arrive over time.  Worker threads, each
owning a build context (its own HIP stream), take every pending job that shares
the oldest pending job's (leaf, bucket) and build them together with one
rs_build_batch call; builds of different workers overlap on the device.  A job
is admitted only while the estimated HBM footprint of everything admitted stays
under the budget; it stays queued otherwise.  Each job records arrival, start and
completion time; its MPHF is the same as a standalone build of its keys
(rs_build_batch builds every set independently).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import Context, GPURecSplit, build_batch
from .keys import splitmix64_keys

# HBM bytes per key of one build: input keys 16 B, bucket-sorted copy 8 B, plus
# node descriptors, seeds and search state (~1 node per 4-8 keys, 58 B each).
BYTES_PER_KEY = 40


@dataclass
class Job:
    job_id: int
    keys: np.ndarray          # (n, 2) uint64
    leaf: int
    bucket: int
    arrival: float            # seconds on the server clock
    start: float = -1.0
    done: float = -1.0
    result: GPURecSplit | None = None
    error: str | None = None

    @property
    def n(self) -> int:
        return int(self.keys.shape[0])

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class ServerStats:
    jobs: int = 0
    keys: int = 0
    batches: int = 0
    wall_s: float = 0.0
    latencies: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"jobs": self.jobs, "keys": self.keys, "batches": self.batches, "wall_s": self.wall_s,
                "objects_per_s": self.keys / self.wall_s if self.wall_s else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_latency_s": float(lat.mean())}


class Server:
    """Continuous-batching construction server over `workers` concurrent contexts."""

    def __init__(self, workers: int = 2, hbm_budget_bytes: int = 64 << 30, max_batch_keys: int = 50_000_000,
                 device: int = 0, time_limit_s: float | None = None, on_batch=None):
        """time_limit_s bounds each batch build (its jobs fail with the time-limit
        error past it); on_batch(jobs) is called after every finished batch."""
        import torch

        self.torch = torch
        self.device = device
        self.on_batch = on_batch
        self.ctxs = [Context(device, time_limit_s=time_limit_s) for _ in range(workers)]
        self.budget = hbm_budget_bytes
        self.max_batch_keys = max_batch_keys
        self.pending: list[Job] = []
        self.admitted_bytes = 0
        self.cv = threading.Condition()
        self.closed = False
        self.t0 = time.perf_counter()
        self.stats = ServerStats()
        self.threads = [threading.Thread(target=self._worker, args=(c,), daemon=True) for c in self.ctxs]
        for t in self.threads:
            t.start()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, job: Job):
        with self.cv:
            self.pending.append(job)
            self.cv.notify_all()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        self.stats.wall_s = self.now()

    def _take(self) -> list[Job] | None:
        """Called with the lock held: the next admissible batch, or None."""
        if not self.pending:
            return None
        head = self.pending[0]
        group, keys, need = [], 0, 0
        for j in self.pending:
            if (j.leaf, j.bucket) != (head.leaf, head.bucket):
                continue
            b = j.n * BYTES_PER_KEY
            if group and (keys + j.n > self.max_batch_keys or self.admitted_bytes + need + b > self.budget):
                break
            if not group and self.admitted_bytes + b > self.budget and self.admitted_bytes > 0:
                return None  # wait for memory to free up
            group.append(j)
            keys += j.n
            need += b
        for j in group:
            self.pending.remove(j)
        self.admitted_bytes += need
        return group

    def _worker(self, ctx: Context):
        torch = self.torch
        while True:
            with self.cv:
                batch = None
                while batch is None:
                    batch = self._take()
                    if batch is None:
                        if self.closed and not self.pending:
                            return
                        self.cv.wait(timeout=0.05)
            start = self.now()
            for j in batch:
                j.start = start
            try:
                keys = np.concatenate([j.keys for j in batch])
                off = np.concatenate([[0], np.cumsum([j.n for j in batch])]).astype(np.uint64)
                d = torch.from_numpy(keys.view(np.int64)).to(f"cuda:{self.device}")
                results, _ = build_batch(d, off, batch[0].bucket, leaf_size=batch[0].leaf, ctx=ctx)
                del d
                for j, r in zip(batch, results):
                    j.result = r
            except Exception as e:  # the whole batch fails together
                for j in batch:
                    j.error = str(e)
            done = self.now()
            with self.cv:
                for j in batch:
                    j.done = done
                    self.stats.jobs += 1
                    self.stats.keys += j.n
                    self.stats.latencies.append(j.latency)
                self.stats.batches += 1
                self.admitted_bytes -= sum(j.n * BYTES_PER_KEY for j in batch)
                self.cv.notify_all()
            if self.on_batch is not None:
                self.on_batch(batch)


def poisson_stream(num_jobs: int, rate_per_s: float, seed: int, size_lo: int, size_hi: int,
                   leaf_weights: dict[int, float], bucket_of_leaf, size_cap_of_leaf=None):
    """Synthetic job stream: Poisson arrivals, log-uniform sizes, leaf sizes drawn
    with the given weights; keys are splitmix64 draws (first, second)."""
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


def run_stream(stream, workers: int = 2, hbm_budget_bytes: int = 64 << 30, time_limit_s: float | None = None,
               on_batch=None) -> tuple[list[Job], dict]:
    """Replays a stream in real time against a Server and returns (jobs, summary).
    Keys are generated before the server clock starts."""
    prepared = [(t_arr, splitmix64_keys(kseed, n), leaf, bucket) for (t_arr, n, leaf, bucket, kseed) in stream]
    srv = Server(workers=workers, hbm_budget_bytes=hbm_budget_bytes, time_limit_s=time_limit_s, on_batch=on_batch)
    jobs = []
    for i, (t_arr, keys, leaf, bucket) in enumerate(prepared):
        delay = t_arr - srv.now()
        if delay > 0:
            time.sleep(delay)
        j = Job(i, keys, leaf, bucket, arrival=srv.now())
        jobs.append(j)
        srv.submit(j)
    srv.close()
    return jobs, srv.stats.summary()
