"""Build plans (rs_plan_*): a whole construction queued on a stream with no
allocation and no host synchronisation.  The plan's bytes equal the ordinary
build's (and so the oracle's), one plan serves many key sets, errors are reported
and leave the plan usable, and the queued build can be captured into a graph and
replayed for new keys."""
import sys

import numpy as np
import pytest

from oracle_lib import OracleMPHF, splitmix_keys

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd as gra
    return torch, gra


def _dev(torch, keys):
    return torch.from_numpy(np.ascontiguousarray(keys).view(np.int64)).to("cuda:0")


PLAN_CASES = [
    (8, 100, 1), (8, 100, 7), (8, 100, 1000), (8, 100, 30000), (12, 100, 20000), (16, 2000, 20000),
    (16, 2000, 123), (18, 50, 5000), (2, 7, 2000), (5, 1009, 6000), (24, 10, 100), (2, 2100, 2100),
]


@pytest.mark.parametrize("leaf,bucket,n", PLAN_CASES)
def test_plan_bytes_match_build(env, leaf, bucket, n):
    torch, gra = env
    _log(f"plan leaf={leaf} bucket={bucket} n={n}")
    keys = splitmix_keys(900 + leaf * 31 + n, n)
    d = _dev(torch, keys)
    plan = gra.Plan(n, bucket, leaf)
    got = plan.build(d)
    assert got == gra.GPURecSplit(d, bucket, leaf_size=leaf).serialize()
    if n <= 30000:
        assert got == OracleMPHF.build(keys, leaf, bucket, threads=8).serialize()
    plan.close()


@pytest.mark.parametrize("rotate,remap64", [(False, False), (True, True)])
def test_plan_modes(env, rotate, remap64):
    torch, gra = env
    n, leaf, bucket = 8000, 10, 100
    _log(f"plan modes rotate={rotate} remap64={remap64}")
    d = _dev(torch, splitmix_keys(4242, n))
    plan = gra.Plan(n, bucket, leaf, rotate=rotate, remap64=remap64)
    assert plan.build(d) == gra.GPURecSplit(d, bucket, leaf_size=leaf, rotate=rotate, remap64=remap64).serialize()
    plan.close()


def test_plan_reused_for_new_keys(env):
    torch, gra = env
    n, leaf, bucket = 25000, 12, 100
    plan = gra.Plan(n, bucket, leaf)
    for seed in (1, 2, 3):
        _log(f"plan reuse seed={seed}")
        d = _dev(torch, splitmix_keys(7000 + seed, n))
        assert plan.build(d) == gra.GPURecSplit(d, bucket, leaf_size=leaf).serialize()
    plan.close()


def test_plan_errors_leave_it_usable(env):
    torch, gra = env
    n, leaf, bucket = 5000, 8, 100
    plan = gra.Plan(n, bucket, leaf)
    keys = splitmix_keys(31, n)
    dup = keys.copy()
    dup[1] = dup[0]  # same bucket, same second
    _log("plan duplicate keys")
    with pytest.raises(gra.RecSplitError) as e:
        plan.build(_dev(torch, dup))
    assert e.value.code == 4
    one = keys.copy()
    one[:, 0] = one[0, 0]  # every key in one bucket of 5000
    _log("plan oversized bucket")
    with pytest.raises(gra.RecSplitError) as e:
        plan.build(_dev(torch, one))
    assert e.value.code == 3
    _log("plan after errors")
    d = _dev(torch, keys)
    assert plan.build(d) == gra.GPURecSplit(d, bucket, leaf_size=leaf).serialize()
    plan.close()


def test_plan_captured_in_a_graph_and_replayed(env):
    """The plan's build recorded once by stream capture; each replay builds the keys
    then in the captured buffer.  Replays alternate with eager builds and every
    replay is checked."""
    torch, gra = env
    n, leaf, bucket = 30000, 8, 100
    plan = gra.Plan(n, bucket, leaf)
    static = _dev(torch, splitmix_keys(555, n))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    _log("plan capture")
    with torch.cuda.graph(g):
        plan.enqueue(static)
    for seed in (11, 12, 13, 14):
        _log(f"plan replay seed={seed}")
        d = _dev(torch, splitmix_keys(seed, n))
        static.copy_(d)
        g.replay()
        torch.cuda.synchronize()
        got = plan.result()
        assert got == gra.GPURecSplit(d, bucket, leaf_size=leaf).serialize()
        assert got == OracleMPHF.build(d.cpu().numpy().view(np.uint64), leaf, bucket, threads=8).serialize()
    del g
    plan.close()
