"""Served construction through the library's server (rs_server_*): jobs of
different sizes and leaf sizes (8, 12, 16, 24) arriving over time are built with
continuous batching; every job's MPHF is byte-identical to its standalone build,
and to the oracle's (directly for a sample, through the stored digest of the
oracle's l = 24 build for the l = 24 reference job); a job with duplicate keys
fails alone without failing the jobs batched with it."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OracleMPHF

pytestmark = pytest.mark.gpu

L24 = json.loads((Path(__file__).resolve().parent / "golden" / "headline_mphf.json").read_text())[
    "plain_l24_b2000_n8000"]


@pytest.fixture(scope="module")
def gra():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd
    return gpurecsplit_amd


def test_served_results_equal_standalone_and_oracle(gra):
    from gpurecsplit_amd.keys import splitmix64_keys
    from gpurecsplit_amd.serve import poisson_stream, run_stream

    stream = poisson_stream(28, rate_per_s=100.0, seed=4, size_lo=500, size_hi=30_000,
                            leaf_weights={8: 0.3, 12: 0.3, 16: 0.3, 24: 0.1},
                            bucket_of_leaf=lambda l: 2000 if l >= 16 else 100,
                            size_cap_of_leaf=lambda l: 3000 if l == 24 else 10**9)
    # job 5: the l = 24 reference key set (digest of the oracle's bytes stored);
    # job 9: an l = 8 job with a duplicated key
    stream[5] = (stream[5][0], L24["n"], 24, 2000, -1)
    stream[9] = (stream[9][0], 4000, 8, 100, -2)

    def keys_of(job):
        if job[4] == -1:
            return splitmix64_keys(L24["seed"], L24["n"])
        if job[4] == -2:
            k = splitmix64_keys(99, job[1])
            k[100] = k[200]
            return k
        return splitmix64_keys(job[4], job[1])

    leaves = {s[2] for s in stream}
    assert 24 in leaves and len(leaves) >= 3
    print(f"stream: {len(stream)} jobs, leaves {sorted(leaves)}", flush=True)
    jobs, summ = run_stream(stream, slots=3, keys_of=keys_of)
    print(f"served: {summ}", flush=True)
    failed = [j for j in jobs if j.error]
    assert [j.job_id for j in failed] == [jobs[9].job_id], [j.error for j in failed]
    assert "duplicate" in failed[0].error
    assert summ["server"]["batches"] < len(jobs)  # arrivals during builds were batched
    for i, j in enumerate(jobs):
        if i == 9:
            continue
        got = j.result.serialize()
        alone = gra.GPURecSplit(j.keys, j.bucket, leaf_size=j.leaf).serialize()
        assert got == alone, f"job {i} (l={j.leaf}, n={j.n}, batch of {j.batch_jobs}): mix != alone"
        if i == 5:
            assert hashlib.sha256(got).hexdigest() == L24["sha256"]
        elif j.leaf != 24 and i % 4 == 0:
            assert got == OracleMPHF.build(j.keys, j.leaf, j.bucket, threads=8).serialize(), f"job {i}"
        assert j.arrival <= j.start <= j.done
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0


def test_server_poll_wait_and_budget(gra):
    """rs_poll / rs_wait through the C ABI; a budget too small for two jobs' keys
    at once still finishes every job (the second waits for the first's memory)."""
    from gpurecsplit_amd.keys import splitmix64_keys
    from gpurecsplit_amd.serve import Server

    keys = [splitmix64_keys(70 + i, 2_000_000) for i in range(3)]  # 32 MB of keys each
    with Server(slots=2, hbm_budget_bytes=400 << 20) as srv:
        ids = [srv.submit(k, 12, 100) for k in keys]
        st = srv.poll(ids[0])
        assert st["state"] in ("queued", "running", "done") and st["n"] == 2_000_000
        res = [srv.wait(i) for i in ids]
        stats = srv.stats()
        assert stats["jobs_done"] == 3 and stats["peak_committed_bytes"] <= (400 << 20) + (300 << 20)
        with pytest.raises(gra.RecSplitError):
            srv.wait(ids[0])  # handed out already
    for k, (rs, info) in zip(keys, res):
        assert info["state"] == "done"
        assert rs.serialize() == gra.GPURecSplit(k, 100, leaf_size=12).serialize()
