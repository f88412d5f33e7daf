"""GPU parity: the HIP build's serialized MPHF is byte-identical to the oracle's
on the same keys, every key maps to a distinct index in [0, n), and the GPU
query kernel agrees with the host query.  Runs through the C ABI."""
import sys

import numpy as np
import pytest

from oracle_lib import OracleMPHF, splitmix_keys

pytestmark = pytest.mark.gpu

gra = pytest.importorskip("gpurecsplit_amd")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _parity(keys, leaf, bucket, rotate=True, threads=8):
    _log(f"parity leaf={leaf} bucket={bucket} n={keys.shape[0]} rotate={rotate}")
    rs = gra.GPURecSplit(keys, bucket, leaf_size=leaf, rotate=rotate)
    ref = OracleMPHF.build(keys, leaf, bucket, rotate=rotate, threads=threads)
    got, exp = rs.serialize(), ref.serialize()
    if got != exp:
        # locate the first differing seed for a readable failure
        g = OracleMPHF.load(got, leaf, rotate)
        gs, es = g.seeds(), ref.seeds()
        k = next((i for i in range(min(len(gs), len(es))) if gs[i] != es[i]), None)
        pytest.fail(f"bytes differ (len {len(got)} vs {len(exp)}); first differing seed index {k}: "
                    f"{gs[k] if k is not None else None} vs {es[k] if k is not None else None}")
    q = rs.query(keys)
    assert np.array_equal(np.sort(q), np.arange(keys.shape[0], dtype=np.uint64))
    return rs


CASES = [
    (2, 7, 2000), (3, 4, 1000), (4, 25, 3000), (5, 10, 5000), (6, 1009, 6000), (7, 100, 10000),
    (8, 100, 50000), (8, 2, 300), (8, 2000, 20000), (9, 231, 8000), (10, 483, 10000), (11, 75, 5000),
    (12, 100, 30000), (13, 1300, 8000), (14, 231, 4000), (16, 2000, 10000), (16, 100, 5000),
    (18, 50, 3000), (20, 100, 500), (24, 10, 100), (8, 5, 1), (8, 100, 10), (16, 2000, 123),
    # leaf 2: buckets of 2049 or more keys have more tree nodes than the memo's
    # 11-bit node field holds
    (2, 2100, 2100), (2, 2000, 20000), (3, 2900, 2900),
]


@pytest.mark.parametrize("leaf,bucket,n", CASES)
def test_bytes_match_oracle(torch_cuda, leaf, bucket, n):
    keys = splitmix_keys(leaf * 7919 + bucket + n, n)
    _parity(keys, leaf, bucket)


def _random_shapes(count=14, seed=20261203):
    """Deterministic draws of (leaf, bucket, n, rotate): leaf sizes 2..18, buckets
    from a few keys to 2000, up to 20,000 keys, both leaf modes."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        leaf = int(rng.integers(2, 19))
        bucket = int(np.exp(rng.uniform(np.log(2), np.log(2000))))
        n = int(rng.integers(1, 20001))
        out.append((leaf, max(bucket, 1), n, bool(rng.integers(0, 4) > 0)))
    return out


@pytest.mark.parametrize("leaf,bucket,n,rotate", _random_shapes())
def test_random_shapes_match_oracle(torch_cuda, leaf, bucket, n, rotate):
    """Shapes nobody picked: every launch kind, grid size and kernel instance (byte
    tables where a level's nodes are mostly full-size, increment tables elsewhere)
    has to give the oracle's bytes."""
    keys = splitmix_keys(leaf * 31 + bucket * 7 + n, n)
    _parity(keys, leaf, bucket, rotate=rotate)


@pytest.mark.parametrize("leaf,bucket,n", [(2, 7, 2000), (5, 10, 5000), (8, 100, 20000), (11, 75, 3000)])
def test_bytes_match_oracle_bruteforce_leaves(torch_cuda, leaf, bucket, n):
    keys = splitmix_keys(leaf + 5, n)
    _parity(keys, leaf, bucket, rotate=False)


def test_reference_cpu_config_1m_l8_b100(torch_cuda):
    """n=1,000,000 splitmix64(42) keys, l=8, b=100, rotation: GPU bytes == oracle bytes."""
    keys = splitmix_keys(42, 1_000_000)
    _parity(keys, 8, 100)


def test_device_query_matches_oracle(torch_cuda):
    """k_query on a GPU-built structure returns the oracle's operator() value for
    every key (AbstractParallelRecSplit.hpp:409-466)."""
    torch = torch_cuda
    keys = splitmix_keys(9, 200_000)
    rs = gra.GPURecSplit(keys, 500, leaf_size=10)
    ref = OracleMPHF.build(keys, 10, 500, threads=8)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    out = rs.query_device(d).cpu().numpy().view(np.uint64)
    assert np.array_equal(out, ref.query(keys))
    assert np.array_equal(np.sort(out), np.arange(200_000, dtype=np.uint64))


@pytest.mark.parametrize("leaf,bucket,n,rotate", [(8, 100, 50_000, True), (16, 2000, 30_000, True),
                                                  (5, 7, 4000, False), (24, 24, 48, True), (2, 2100, 2100, True)])
def test_device_query_of_deserialized_oracle_bytes(torch_cuda, leaf, bucket, n, rotate):
    """A structure loaded from the oracle's bytes (no build on this context) is
    uploaded on its first device query; results equal the oracle's operator()."""
    torch = torch_cuda
    keys = splitmix_keys(300 + leaf, n)
    ref = OracleMPHF.build(keys, leaf, bucket, rotate=rotate, threads=8)
    rs = gra.GPURecSplit.deserialize(ref.serialize(), leaf, rotate=rotate)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    out = rs.query_device(d).cpu().numpy().view(np.uint64)
    assert np.array_equal(out, ref.query(keys))
    out2 = rs.query_device(d[: n // 2 + 1]).cpu().numpy().view(np.uint64)  # reuses the device copy
    assert np.array_equal(out2, out[: n // 2 + 1])


def test_device_query_of_batch_results(torch_cuda):
    """Handles returned by a batch build have no device copy until their first
    device query."""
    torch = torch_cuda
    sizes = [5000, 17, 30000]
    keys = np.concatenate([splitmix_keys(400 + i, s) for i, s in enumerate(sizes)])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    rss, _ = gra.build_batch(d, off, 100, leaf_size=8)
    for i, rs in enumerate(rss):
        k = keys[off[i]:off[i + 1]]
        out = rs.query_device(d[int(off[i]):int(off[i + 1])]).cpu().numpy().view(np.uint64)
        assert np.array_equal(out, OracleMPHF.build(k, 8, 100).query(k)), f"set {i}"


def test_device_resident_keys_build(torch_cuda):
    torch = torch_cuda
    keys = splitmix_keys(10, 40_000)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    a = gra.GPURecSplit(d, 300, leaf_size=9).serialize()
    b = gra.GPURecSplit(keys, 300, leaf_size=9).serialize()
    assert a == b


def test_repeated_builds_reuse_context(torch_cuda):
    """Buffers are reused across builds; a bigger then a smaller build must both
    match the oracle."""
    for n in (30_000, 2_000, 30_000):
        keys = splitmix_keys(n, n)
        _parity(keys, 8, 150)


def test_duplicate_keys_rejected(torch_cuda):
    keys = splitmix_keys(11, 1000)
    keys[10] = keys[20]
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit(keys, 50, leaf_size=8)
    assert e.value.code == 4


def test_equal_second_in_bucket_rejected(torch_cuda):
    keys = splitmix_keys(12, 1000)
    keys[30, 1] = keys[31, 1]
    keys[30, 0] = keys[31, 0] ^ 1  # same bucket (differ in the lowest bit of first)
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit(keys, 50, leaf_size=8)
    assert e.value.code == 4


def test_oversized_bucket_rejected(torch_cuda):
    keys = splitmix_keys(13, 3500)
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit(keys, 5000, leaf_size=8)
    assert e.value.code == 3


def test_stats_and_evals(torch_cuda):
    keys = splitmix_keys(14, 50_000)
    rs = gra.GPURecSplit(keys, 2000, leaf_size=12)
    st = rs.stats()
    ref = OracleMPHF.build(keys, 12, 2000, threads=8)
    assert sum(st["evals"].values()) == ref.evals
    assert st["num_nodes"] == ref.nseeds


def test_batch_each_set_matches_oracle(torch_cuda):
    """rs_build_batch: every set's bytes equal the oracle's build of that set alone."""
    torch = torch_cuda
    rng = np.random.default_rng(5)
    sizes = [1, 2, 10, 123, 999, 5000, 20000, 1, 3000, 777]
    keys = np.concatenate([splitmix_keys(1000 + i, s) for i, s in enumerate(sizes)])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    rss, st = gra.build_batch(d, off, 100, leaf_size=12)
    assert len(rss) == len(sizes)
    for i, rs in enumerate(rss):
        k = keys[off[i]:off[i + 1]]
        assert rs.serialize() == OracleMPHF.build(k, 12, 100).serialize(), f"set {i} (n={sizes[i]})"
    assert st["num_buckets"] == sum(max(1, -(-s // 100)) for s in sizes)


def test_batch_many_uneven_sets(torch_cuda):
    """512 sets with log-uniform sizes (the batched workload's shape, scaled down)."""
    torch = torch_cuda
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(100), np.log(20000), size=512)).astype(np.int64)
    keys = splitmix_keys(3, int(sizes.sum()))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    rss, st = gra.build_batch(d, off, 100, leaf_size=12)
    for i in range(0, 512, 37):
        k = keys[off[i]:off[i + 1]]
        assert rss[i].serialize() == OracleMPHF.build(k, 12, 100).serialize(), f"set {i}"
    for i in range(512):
        q = rss[i].query(keys[off[i]:off[i + 1]])
        assert np.array_equal(np.sort(q), np.arange(sizes[i], dtype=np.uint64)), f"set {i}"


def test_time_limit_stops_build_and_context_recovers(torch_cuda):
    """A 1 us limit ends every search wave at its first claim (RS_ERR_TIMEOUT = 10);
    with the limit lifted the same context builds the oracle's bytes."""
    keys = splitmix_keys(15, 200_000)  # l=16, b=2000: tens of ms of search
    ctx = gra.Context(0, time_limit_s=1e-6)
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit(keys, 2000, leaf_size=16, ctx=ctx)
    assert e.value.code == 10
    ctx.set_time_limit(0)
    rs = gra.GPURecSplit(keys, 2000, leaf_size=16, ctx=ctx)
    assert rs.serialize() == OracleMPHF.build(keys, 16, 2000, threads=8).serialize()
    ctx.set_time_limit(600)  # a generous limit changes nothing
    assert gra.GPURecSplit(keys, 2000, leaf_size=16, ctx=ctx).serialize() == rs.serialize()


def test_l24_leaf_seeds_above_2_31(torch_cuda):
    """24-key rotation leaves whose lowest seeds lie above 2^31 and 2^32 (found by
    the oracle): the search must stop at them (wave-uniform 64-bit values keep
    their high word) and return the same bytes."""
    for s in (100, 101, 103):  # oracle seeds 4433910963, 610147800, 2643051985
        keys = splitmix_keys(s, 24)
        ref = OracleMPHF.build(keys, 24, 24, threads=1)
        rs = gra.GPURecSplit(keys, 24, leaf_size=24)
        assert rs.serialize() == ref.serialize(), f"seed {s}"
    assert max(OracleMPHF.build(splitmix_keys(100, 24), 24, 24).seeds()) >= 2**32


def test_wide_counters_use_high_word(torch_cuda):
    """l=20, one 900-key bucket: the upper node has 6 parts (5 counted 8-bit fields,
    40 bits) and the lower nodes 8 parts (7 counted 5-bit fields, 35 bits), so the
    64-bit counters carry fields above bit 31."""
    keys = splitmix_keys(20, 900)
    _parity(keys, 20, 900, threads=1)


R64_CASES = [(2, 7, 2000, True), (5, 10, 5000, False), (8, 100, 50000, True), (8, 100, 20000, False),
             (12, 100, 30000, True), (16, 2000, 10000, True), (18, 50, 3000, True), (24, 10, 100, True)]


@pytest.mark.parametrize("leaf,bucket,n,rotate", R64_CASES)
def test_remap64_bytes_match_oracle(torch_cuda, leaf, bucket, n, rotate):
    """The RecSplitRotate variant (64-bit remap, RecSplitRotate.hpp:115): GPU bytes
    equal the oracle's, and the GPU query kernel agrees with the oracle's query."""
    torch = torch_cuda
    keys = splitmix_keys(leaf * 31 + n, n)
    flags = (1 if rotate else 0) | 2
    _log(f"remap64 leaf={leaf} bucket={bucket} n={n} rotate={rotate}")
    ref = OracleMPHF.build(keys, leaf, bucket, rotate=flags, threads=8)
    rs = gra.GPURecSplit(keys, bucket, leaf_size=leaf, rotate=rotate, remap64=True)
    assert rs.serialize() == ref.serialize()
    d = torch.from_numpy(keys.view(np.int64)).cuda()
    assert np.array_equal(rs.query_device(d).cpu().numpy().view(np.uint64), ref.query(keys))


def test_wave_times_keep_bytes_and_report_shares(torch_cuda, monkeypatch):
    """RS_WAVE_TIMES only records the search waves' start and exit: the bytes equal
    the oracle's, and each searched kind reports a busy share in (0, 1] and a
    non-negative exit tail."""
    monkeypatch.setenv("RS_WAVE_TIMES", "1")
    ctx = gra.Context(0)
    monkeypatch.delenv("RS_WAVE_TIMES")
    keys = splitmix_keys(4242, 20000)
    _log("wave times leaf=12 bucket=100 n=20000")
    rs = gra.GPURecSplit(keys, 100, leaf_size=12, ctx=ctx)
    assert rs.serialize() == OracleMPHF.build(keys, 12, 100, threads=8).serialize()
    st = rs.stats()
    searched = [k for k, ms in st["search_ms"].items() if ms > 0]
    assert searched
    for k in searched:
        assert 0 < st["search_busy"][k] <= 1.0, (k, st["search_busy"][k])
        assert st["search_tail_ms"][k] >= 0
    plain = gra.GPURecSplit(keys, 100, leaf_size=12).stats()
    assert all(v == 0 for v in plain["search_busy"].values())
    rs.close()
    ctx.close()


def _tree_sizes(m, leaf):
    """Node sizes of a bucket's split tree in DFS preorder (children of size 1 carry
    no seed), by the split rules of AbstractParallelRecSplit.hpp:133-148 and the
    fanout-2 split size of SIMDRecSplit.hpp:632."""
    lower = leaf * max(2, -(-(35 * leaf + 50) // 100))
    upper = lower * -(-(21 * leaf + 90) // 100)
    out, stack = [], [m]
    while stack:
        s = stack.pop()
        out.append(s)
        if s <= leaf:
            continue
        if s > upper:
            split = ((s // 2 + upper - 1) // upper) * upper
            kids = [split, s - split]
        else:
            unit = lower if s > lower else leaf
            fan = -(-s // unit)
            kids = [unit] * (fan - 1) + [s - unit * (fan - 1)]
        stack.extend(k for k in reversed(kids) if k > 1)
    return out


def test_seed_counts_from_the_seeds(torch_cuda):
    """rs_build_stats.seeds: x + 1 per split or brute-force node, (x // L + 1) * L per
    rotation leaf, recomputed from the structure's own seeds (decoded by the oracle)
    and node sizes (a Python tree walk over the bucket sizes)."""
    leaf, bucket, n = 12, 100, 20000
    keys = splitmix_keys(77, n)
    rs = gra.GPURecSplit(keys, bucket, leaf_size=leaf)
    st = rs.stats()
    seeds = OracleMPHF.load(rs.serialize(), leaf).seeds()
    nb = -(-n // bucket)
    first = keys[:, 0]
    hi, lo = first >> np.uint64(32), first & np.uint64(0xFFFFFFFF)
    b = (hi * np.uint64(nb) + ((lo * np.uint64(nb)) >> np.uint64(32))) >> np.uint64(32)
    sizes = np.bincount(b.astype(np.int64), minlength=nb)
    expect, i = 0, 0
    for m in sizes:
        if m <= 1:
            continue
        for s in _tree_sizes(int(m), leaf):
            x = int(seeds[i])
            i += 1
            expect += (x // leaf + 1) * leaf if s == leaf else x + 1
    assert i == len(seeds)
    assert sum(st["seeds"].values()) == expect
    assert st["seeds"]["leaf_rotation"] == st["evals"]["leaf_rotation"]


def test_small_levels_run_on_fewer_waves(torch_cuda):
    """Levels of little expected work launch fewer waves than the full grid, and
    the bytes do not depend on the grid (they equal the oracle's)."""
    keys = splitmix_keys(4040, 10_000)
    _log("small levels leaf=8 bucket=100 n=10000")
    rs = gra.GPURecSplit(keys, 100, leaf_size=8)
    st = rs.stats()
    full = st["search_waves"]
    waves = {k: v for k, v in st["kind_waves"].items() if st["nodes"][k]}
    assert waves and all(0 < v <= full for v in waves.values()), (waves, full)
    assert min(waves.values()) < full
    assert rs.serialize() == OracleMPHF.build(keys, 8, 100, threads=8).serialize()
