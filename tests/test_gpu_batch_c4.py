"""Batched build in the shape of the 4096-set workload (l=12, b=100, per-set
sizes log-uniform, size seed 3, keys from one SplitMix64(3) stream generated on
the GPU), with sizes scaled to [1e3, 2e4] so that the host checks stay short:
every set is a bijection onto [0, n_i), and sampled sets are byte-identical to
the oracle's build of that set alone."""
import sys

import numpy as np
import pytest

from oracle_lib import OracleMPHF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd as gra
    from gpurecsplit_amd.keys import splitmix64_keys_device
    return torch, gra, splitmix64_keys_device


def test_4096_uneven_sets_l12_b100(env):
    torch, gra, gen = env
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(1_000), np.log(20_000), size=4096)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    print(f"4096 sets, {n} keys", flush=True)
    d = gen(3, n)
    rss, st = gra.build_batch(d, off, 100, leaf_size=12)
    assert len(rss) == 4096
    assert st["num_buckets"] == int(sum(-(-int(s) // 100) for s in sizes))
    keys = d.cpu().numpy().view(np.uint64)
    del d
    for i in range(4096):
        k = keys[off[i]:off[i + 1]]
        q = rss[i].query(k)
        assert np.array_equal(np.sort(q), np.arange(sizes[i], dtype=np.uint64)), f"set {i}"
    for i in range(0, 4096, 128):
        k = keys[off[i]:off[i + 1]]
        assert rss[i].serialize() == OracleMPHF.build(k, 12, 100, threads=8).serialize(), f"set {i}"
    sys.stdout.flush()


def test_batch_set_above_16384_buckets_uses_global_atomics(env):
    """A set of 1,700,000 keys at b = 100 has 17,000 buckets, more than the LDS
    histogram and slot reservation hold: its bucket counts and slots go through
    global atomics, the small sets' through LDS.  Every set's bytes equal the
    oracle's build of that set alone."""
    torch, gra, gen = env
    sizes = [1_700_000, 5000, 300]
    keys = np.concatenate([gen(500 + i, s).cpu().numpy().view(np.uint64) for i, s in enumerate(sizes)])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    print(f"sets {sizes}", flush=True)
    rss, st = gra.build_batch(d, off, 100, leaf_size=12)
    assert st["num_buckets"] == 17000 + 50 + 3
    for i, rs in enumerate(rss):
        k = keys[off[i]:off[i + 1]]
        assert rs.serialize() == OracleMPHF.build(k, 12, 100, threads=8).serialize(), f"set {i}"


def test_batch_more_than_4M_buckets_device_scan(env):
    """Two sets at b = 1 with 4.5M buckets in total: every bucket scan has more
    than 2048 blocks of 2048 items, so the block sums are scanned in two device
    levels.  Bytes equal the oracle's."""
    torch, gra, gen = env
    sizes = [2_200_000, 2_300_000]
    keys = np.concatenate([gen(n, n).cpu().numpy().view(np.uint64) for n in sizes])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    print(f"sets {sizes} at b=1", flush=True)
    rss, st = gra.build_batch(d, off, 1, leaf_size=8)
    assert st["num_buckets"] == sum(sizes)
    for i, rs in enumerate(rss):
        k = keys[off[i]:off[i + 1]]
        assert rs.serialize() == OracleMPHF.build(k, 8, 1, threads=8).serialize(), f"set {i}"
    one = gra.GPURecSplit(d[: sizes[0] + sizes[1]], 1, leaf_size=8)  # plain build, 4.5M buckets
    k = keys[: sizes[0] + sizes[1]]
    assert one.serialize() == OracleMPHF.build(k, 8, 1, threads=8).serialize()


def test_batch_of_34M_keys_uses_large_tiles(env):
    """34.4M keys in one batch (more than two 64K-key tiles per CU, so the
    histogram and scatter take 64K-key tiles): one set of 2M keys above the LDS
    bucket limit and 108 sets of 300,000 keys below it.  Every set is a bijection
    (device queries) and two sets are byte-identical to the oracle's builds."""
    torch, gra, gen = env
    sizes = [2_000_000] + [300_000] * 108
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    print(f"{len(sizes)} sets, {n} keys", flush=True)
    d = gen(77, n)
    rss, st = gra.build_batch(d, off, 100, leaf_size=12)
    assert st["num_buckets"] == 20000 + 108 * 3000
    for i, rs in enumerate(rss):
        a, b = int(off[i]), int(off[i + 1])
        idx = rs.query_device(d[a:b])
        assert torch.equal(torch.sort(idx).values, torch.arange(b - a, device=idx.device)), f"set {i}"
    for i in (1, 108):
        k = d[int(off[i]):int(off[i + 1])].cpu().numpy().view(np.uint64)
        assert rss[i].serialize() == OracleMPHF.build(k, 12, 100, threads=8).serialize(), f"set {i}"
