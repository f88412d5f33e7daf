"""Served construction: jobs of different sizes and leaf sizes arriving over
time, built with continuous batching on concurrent contexts; every job's MPHF is
byte-identical to its standalone build and to the oracle's, alone or in the mix."""
import numpy as np
import pytest

from oracle_lib import OracleMPHF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gra():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd
    return gpurecsplit_amd


def test_served_results_equal_standalone(gra):
    from gpurecsplit_amd.serve import poisson_stream, run_stream

    stream = poisson_stream(24, rate_per_s=200.0, seed=4, size_lo=500, size_hi=30_000,
                            leaf_weights={8: 0.4, 12: 0.3, 16: 0.3},
                            bucket_of_leaf=lambda l: 2000 if l >= 16 else 100)
    jobs, summ = run_stream(stream, workers=2)
    assert summ["jobs"] == 24 and not any(j.error for j in jobs), [j.error for j in jobs if j.error]
    assert summ["batches"] < 24  # arrivals during a build were batched together
    for j in jobs:
        got = j.result.serialize()
        alone = gra.GPURecSplit(j.keys, j.bucket, leaf_size=j.leaf).serialize()
        assert got == alone, f"job {j.job_id}: mix != alone"
        if j.job_id % 4 == 0:
            assert got == OracleMPHF.build(j.keys, j.leaf, j.bucket, threads=8).serialize(), f"job {j.job_id}"
        assert j.arrival <= j.start <= j.done
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0
