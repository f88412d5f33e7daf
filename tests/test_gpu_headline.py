"""Headline workloads at their stated sizes, byte-checked against the oracle.

The oracle's MPHF of each workload is stored as a SHA-256 of its serialized
bytes in tests/golden/headline_mphf.json (written by
tests/golden/make_headline_fixtures.py; the 5M-key builds take minutes of CPU
time).  The GPU build must reproduce those bytes exactly, map every key to a
distinct index through the GPU query kernel, and count the same algorithmic
evaluations as the oracle.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURES = json.loads((Path(__file__).resolve().parent / "golden" / "headline_mphf.json").read_text())


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd as gra
    from gpurecsplit_amd.keys import splitmix64_keys_device
    return torch, gra, splitmix64_keys_device


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_headline_bytes_match_oracle(env, name):
    torch, gra, gen = env
    w = FIXTURES[name]
    print(f"{name}: n={w['n']} seed={w['seed']} l={w['leaf']} b={w['bucket']}", flush=True)
    sys.stdout.flush()
    d = gen(w["seed"], w["n"])
    from gpurecsplit_amd.keys import splitmix64_keys
    assert np.array_equal(d[:1000].cpu().numpy().view(np.uint64), splitmix64_keys(w["seed"], 1000))
    rs = gra.GPURecSplit(d, w["bucket"], leaf_size=w["leaf"], rotate=w["rotate"])
    blob = rs.serialize()
    assert len(blob) == w["bytes"]
    assert hashlib.sha256(blob).hexdigest() == w["sha256"]
    st = rs.stats()
    assert st["num_nodes"] == w["nseeds"]
    assert sum(st["evals"].values()) == w["evals"]
    idx = rs.query_device(d)
    assert bool(torch.equal(torch.sort(idx).values, torch.arange(w["n"], device=idx.device)))
    print(f"{name}: ok, build {st['total_ms']:.1f} ms", flush=True)
