"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle (a wrong constant,
shift, seed offset, rotation direction, carry, split size or bit layout) fails at
least one of them; none of them re-types the oracle's formulas.
"""
import itertools
import json
import random
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import START_SEED, OracleError, OracleMPHF, aggr_bounds, golomb_memo, lib, ptr32, ptr64, splitmix_keys

GOLDEN = Path(__file__).resolve().parent / "golden"
M64 = (1 << 64) - 1


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & M64
    return h


# --------------------------------------------------------------------------- memo


@pytest.mark.parametrize("leaf", range(2, 25))
def test_memo_matches_reference_table(leaf):
    """Golomb-Rice memo regenerated from the formula == the reference's committed
    table (golombRiceMemo.hpp), via the checksum fixture."""
    fx = json.loads((GOLDEN / "golomb_memo.json").read_text())[str(leaf)]
    memo = golomb_memo(leaf)
    for m, v in fx["samples"].items():
        assert int(memo[int(m)]) == v, (leaf, m)
    assert f"{fnv1a64(memo.astype('<u4').tobytes()):016x}" == fx["fnv1a64"]


def test_aggr_bounds_table():
    """lower/upper aggregation bounds for the leaf sizes used by the workloads
    (values derived by hand from AbstractParallelRecSplit.hpp:133-135)."""
    assert aggr_bounds(8) == (32, 96)
    assert aggr_bounds(12) == (60, 240)
    assert aggr_bounds(16) == (112, 560)
    assert aggr_bounds(18) == (126, 630)
    assert aggr_bounds(24) == (216, 1296)
    assert aggr_bounds(2) == (4, 8)


# --------------------------------------------------------------------------- hashes


def test_remix32_is_murmur3_fmix32():
    """Published MurmurHash3_x86_32 vectors for inputs whose body is empty or all
    zero reduce to fmix32 of a known word: ("",1)->0x514E28B7,
    ("",0xffffffff)->0x81F16F39, ("\\0\\0\\0\\0",0)->0x2362F9DE."""
    assert lib.orc_remix32(0) == 0
    assert lib.orc_remix32(1) == 0x514E28B7
    assert lib.orc_remix32(0xFFFFFFFF) == 0x81F16F39
    assert lib.orc_remix32(0xE6546B60) == 0x2362F9DE


def test_splitmix64_known_outputs():
    """SplitMix64 from state 0 (Vigna's reference generator): the first outputs
    are 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F."""
    k = splitmix_keys(0, 2)
    assert [int(v) for v in k.reshape(-1)[:3]] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_remap_range_and_lemire_form():
    """remap(x, n) lies in [0, n) and is the 16-bit Lemire reduction of the
    folded hash: n=1 -> 0, and n=65536 returns the 16-bit hash itself."""
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 2**63, size=200, dtype=np.uint64):
        x = int(x)
        h16 = lib.orc_remix32(((x >> 32) ^ x) & 0xFFFFFFFF) & 0xFFFF
        assert lib.orc_remap(x, 1) == 0
        assert lib.orc_remap(x, 65536) == h16
        for n in (2, 3, 7, 100, 2999):
            assert lib.orc_remap(x, n) == (h16 * n) >> 16


def test_remap128_bucket():
    assert lib.orc_remap128(0, 100) == 0
    assert lib.orc_remap128(M64, 100) == 99
    assert lib.orc_remap128(1 << 63, 10) == 5


# --------------------------------------------------------------------------- minimality


def _keys(seed, m):
    return splitmix_keys(seed, m)[:, 1].copy()


def _remap(x, n):
    return lib.orc_remap(x & M64, n)


def test_higher_level_minimal():
    """Lowest x with #{remap(k+S[level]+x, m) < split} == split, by brute force."""
    for trial in range(40):
        m, split, level = random.Random(trial).choice([(6, 3), (8, 4), (9, 4), (12, 8), (5, 2)]) + (trial % 5,)
        keys = _keys(100 + trial, m)
        ev = np.zeros(1, dtype=np.uint64)
        got = lib.orc_search_higher(ptr64(keys), m, split, level, ptr64(ev))
        x = 0
        while sum(_remap(int(k) + START_SEED[level] + x, m) < split for k in keys) != split:
            x += 1
        assert got == x
        assert int(ev[0]) == (x + 1) * m


def test_aggr_level_minimal():
    """Lowest x whose part counts floor(remap/SPLIT) are exactly
    (SPLIT, ..., SPLIT, m - (f-1) SPLIT), by brute force with plain counting."""
    for trial in range(40):
        m, split = random.Random(trial).choice([(6, 2), (7, 3), (9, 3), (10, 4), (8, 4), (11, 2)])
        seed = START_SEED[17 + trial % 2]
        keys = _keys(200 + trial, m)
        got = lib.orc_search_aggr(ptr64(keys), m, split, seed, None)
        f = (m + split - 1) // split
        target = [split] * (f - 1) + [m - (f - 1) * split]
        x = 0
        while True:
            c = [0] * f
            for k in keys:
                c[_remap(int(k) + seed + x, m) // split] += 1
            if c == target:
                break
            x += 1
        assert got == x


def _rotl(v, r, L):
    return ((v << r) | (v >> (L - r))) & ((1 << L) - 1) if r else v


def test_leaf_rotation_minimal():
    """Lowest seed s = base + r (base a multiple of L, 0 <= r < L) such that the
    even keys' mask OR the odd keys' mask rotated left by r covers [0, L)."""
    S = START_SEED[19]
    for trial in range(30):
        L = 3 + trial % 6
        keys = _keys(300 + trial, L)
        got = lib.orc_search_leaf_rot(ptr64(keys), L, None)
        s = 0
        while True:
            base, r = s - s % L, s % L
            ml = mr = 0
            for k in keys:
                bit = 1 << _remap(int(k) + S + base, L)
                if int(k) % 2 == 0:
                    ml |= bit
                else:
                    mr |= bit
            if ml | _rotl(mr, r, L) == (1 << L) - 1:
                break
            s += 1
        assert got == s, (L, trial)


def test_leaf_bruteforce_minimal():
    S = START_SEED[19]
    for trial in range(30):
        m = 2 + trial % 7
        keys = _keys(400 + trial, m)
        got = lib.orc_search_leaf_brute(ptr64(keys), m, None)
        x = 0
        while len({_remap(int(k) + S + x, m) for k in keys}) != m:
            x += 1
        assert got == x


def test_leaf_bruteforce_midstop_lossless():
    """Sizes >= 12 use the midstop prefilter; the result must still be the
    lowest bijective seed (checked with a vectorised numpy scan)."""
    S = START_SEED[19]
    for trial, m in enumerate([12, 13]):
        keys = _keys(500 + trial, m)
        got = int(lib.orc_search_leaf_brute(ptr64(keys), m, None))
        lo = 0
        found = None
        while found is None:
            xs = np.arange(lo, lo + 50000, dtype=np.uint64)
            masks = np.zeros(xs.shape, dtype=np.uint32)
            for k in keys:
                z = xs + np.uint64((int(k) + S) & M64)
                r = _np_remap(z, m)
                masks |= np.uint32(1) << r
            hit = np.nonzero(masks == (1 << m) - 1)[0]
            if hit.size:
                found = lo + int(hit[0])
            lo += 50000
        assert got == found


def _np_fmix32(z):
    z = z.astype(np.uint32)
    z ^= z >> np.uint32(16)
    z = (z.astype(np.uint64) * np.uint64(0x85EBCA6B)).astype(np.uint32)
    z ^= z >> np.uint32(13)
    z = (z.astype(np.uint64) * np.uint64(0xC2B2AE35)).astype(np.uint32)
    z ^= z >> np.uint32(16)
    return z


def _np_remap(z, n):
    folded = ((z >> np.uint64(32)) ^ (z & np.uint64(0xFFFFFFFF))).astype(np.uint32)
    h = _np_fmix32(folded) & np.uint32(0xFFFF)
    return ((h.astype(np.uint64) * np.uint64(n)) >> np.uint64(16)).astype(np.uint32)


def test_numpy_remap_agrees():
    z = splitmix_keys(9, 1000)[:, 0].copy()
    got = _np_remap(z, 1234)
    exp = np.array([lib.orc_remap(int(v), 1234) for v in z], dtype=np.uint32)
    assert np.array_equal(got, exp)


# --------------------------------------------------------------------------- packed counters


def _packed(parts, m, split):
    """Reference two-word packed form; also asserts the oracle's one-word search
    form agrees with it."""
    a = np.ascontiguousarray(parts, dtype=np.uint32)
    ref = bool(lib.orc_packed_aggr_predicate(ptr32(a), m, split))
    assert bool(lib.orc_search_form_aggr_predicate(ptr32(a), m, split)) == ref
    return ref


def _exact(counts, m, split):
    f = (m + split - 1) // split
    return list(counts) == [split] * (f - 1) + [m - (f - 1) * split]


def test_packed_counters_exhaustive_small():
    """The reference's 8-bit packed counters accept exactly the exact-count
    vectors (SIMDRecSplit.hpp:342-388), here over every count vector."""
    for m, split in [(10, 3), (9, 2), (12, 4), (11, 5), (14, 2)]:
        f = (m + split - 1) // split
        for cuts in itertools.combinations(range(m + f - 1), f - 1):
            counts, prev = [], -1
            for c in cuts + (m + f - 1,):
                counts.append(c - prev - 1)
                prev = c
            parts = [p for p, c in enumerate(counts) for _ in range(c)]
            assert _packed(parts, m, split) == _exact(counts, m, split), (m, split, counts)


def test_packed_counters_carry_cases():
    """Count vectors built so that 8-bit fields overflow and carry into the next
    field; every one of them must be rejected (SPLIT < 255 exactness)."""
    rng = random.Random(7)
    for split in (100, 200, 254):
        for f in range(2, 10):
            m = (f - 1) * split + rng.randint(1, split)
            t = [split] * (f - 1) + [m - (f - 1) * split]
            for i in range(f - 1):
                c = list(t)
                c[i] += 256
                c[i + 1] -= 1
                c[-1] -= 255
                if min(c) < 0:
                    continue
                parts = [p for p, n in enumerate(c) for _ in range(n)]
                assert not _packed(parts, sum(c), split)
            parts = [p for p, n in enumerate(t) for _ in range(n)]
            assert _packed(parts, m, split)
            for _ in range(50):  # random near-misses
                c = list(t)
                a, b = rng.sample(range(f), 2)
                d = rng.choice([1, 255, 256, 257])
                if c[a] >= d:
                    c[a] -= d
                    c[b] += d
                    parts = [p for p, n in enumerate(c) for _ in range(n)]
                    assert _packed(parts, m, split) == _exact(c, m, split)


# --------------------------------------------------------------------------- 64-bit remap


def test_remap64_is_48_bit_lemire_of_remix():
    """The RecSplitRotate remap (AbstractParallelRecSplit.hpp:327-332): the low 48
    bits of the 64-bit finalizer (pinned by the SplitMix64 outputs above), scaled
    to [0, n): n = 1 gives 0, n = 2^16 gives bits 32..47 of the finalizer, and any
    n matches 128-bit arithmetic."""
    rng = np.random.default_rng(2)
    for x in rng.integers(0, 2**63, size=200, dtype=np.uint64):
        x = int(x) * 2 + 1
        r = lib.orc_remix(x & M64)
        assert lib.orc_remap64(x & M64, 1) == 0
        assert lib.orc_remap64(x & M64, 65536) == (r >> 32) & 0xFFFF
        for n in (2, 3, 24, 100, 2999):
            assert lib.orc_remap64(x & M64, n) == ((r & ((1 << 48) - 1)) * n) >> 48


def _remap64(x, n):
    return lib.orc_remap64(x & M64, n)


def test_r64_searches_minimal():
    """Every node search in 64-bit-remap mode returns the brute-force lowest seed
    (the same predicates as above, with the 64-bit remap)."""
    S19 = START_SEED[19]
    for trial in range(24):
        m, split, level = random.Random(trial).choice([(6, 3), (8, 4), (12, 8), (5, 2)]) + (trial % 4,)
        keys = _keys(500 + trial, m)
        got = lib.orc_search_higher_r64(ptr64(keys), m, split, level, None)
        x = 0
        while sum(_remap64(int(k) + START_SEED[level] + x, m) < split for k in keys) != split:
            x += 1
        assert got == x
        m, split = random.Random(trial).choice([(6, 2), (9, 3), (10, 4), (11, 2)])
        seed = START_SEED[17 + trial % 2]
        keys = _keys(600 + trial, m)
        got = lib.orc_search_aggr_r64(ptr64(keys), m, split, seed, None)
        f = (m + split - 1) // split
        target = [split] * (f - 1) + [m - (f - 1) * split]
        x = 0
        while True:
            c = [0] * f
            for k in keys:
                c[_remap64(int(k) + seed + x, m) // split] += 1
            if c == target:
                break
            x += 1
        assert got == x
        m = 2 + trial % 7
        keys = _keys(700 + trial, m)
        got = lib.orc_search_leaf_brute_r64(ptr64(keys), m, None)
        x = 0
        while len({_remap64(int(k) + S19 + x, m) for k in keys}) != m:
            x += 1
        assert got == x
        L = 3 + trial % 6
        keys = _keys(800 + trial, L)
        got = lib.orc_search_leaf_rot_r64(ptr64(keys), L, None)
        s = 0
        while True:
            base, r = s - s % L, s % L
            ml = mr = 0
            for k in keys:
                bit = 1 << _remap64(int(k) + S19 + base, L)
                if int(k) % 2 == 0:
                    ml |= bit
                else:
                    mr |= bit
            if ml | _rotl(mr, r, L) == (1 << L) - 1:
                break
            s += 1
        assert got == s, (L, trial)


@pytest.mark.parametrize("leaf,bucket,n", [(2, 7, 1000), (5, 10, 3000), (8, 100, 10000), (12, 100, 8000),
                                           (16, 2000, 4000), (18, 50, 1500)])
def test_r64_mphf_bijection_and_differs(leaf, bucket, n):
    from oracle_lib import FLAG_REMAP64, FLAG_ROTATE
    keys = splitmix_keys(leaf * 77 + bucket, n)
    rs = OracleMPHF.build(keys, leaf, bucket, rotate=FLAG_ROTATE | FLAG_REMAP64)
    assert rs.bijection_errors(keys) == 0
    back = OracleMPHF.load(rs.serialize(), leaf, FLAG_ROTATE | FLAG_REMAP64)
    assert np.array_equal(back.query(keys), rs.query(keys))
    if n > 100:
        assert rs.serialize() != OracleMPHF.build(keys, leaf, bucket, rotate=True).serialize()
    if leaf <= 8:
        rs2 = OracleMPHF.build(keys, leaf, bucket, rotate=FLAG_REMAP64)
        assert rs2.bijection_errors(keys) == 0


# --------------------------------------------------------------------------- whole MPHF


@pytest.mark.parametrize(
    "leaf,bucket,n",
    [(2, 7, 1000), (3, 4, 123), (5, 10, 5000), (8, 100, 10000), (8, 2, 500), (12, 100, 20000),
     (14, 231, 3000), (16, 2000, 4000), (18, 50, 2000), (6, 1009, 4000), (4, 25, 1), (24, 10, 48)],
)
def test_mphf_bijection(leaf, bucket, n):
    keys = splitmix_keys(leaf * 1000 + bucket, n)
    rs = OracleMPHF.build(keys, leaf, bucket, rotate=True)
    assert rs.bijection_errors(keys) == 0
    if leaf <= 12:  # brute-force full leaves grow as e^l: keep to small leaves
        rs2 = OracleMPHF.build(keys, leaf, bucket, rotate=False)
        assert rs2.bijection_errors(keys) == 0


def test_thread_count_invariance():
    keys = splitmix_keys(77, 30000)
    a = OracleMPHF.build(keys, 10, 100, threads=1).serialize()
    b = OracleMPHF.build(keys, 10, 100, threads=5).serialize()
    assert a == b


def test_key_order_invariance():
    keys = splitmix_keys(78, 20000)
    perm = np.random.default_rng(0).permutation(keys.shape[0])
    a = OracleMPHF.build(keys, 9, 150).serialize()
    b = OracleMPHF.build(keys[perm], 9, 150).serialize()
    assert a == b


def _parse(data: bytes):
    """Splits the serialized stream (AbstractParallelRecSplit.hpp:480-488,
    DoubleEF.hpp:100-113) into its fields."""
    off = 0

    def u64():
        nonlocal off
        v = struct.unpack_from("<Q", data, off)[0]
        off += 8
        return v

    def vec():
        nonlocal off
        n = u64()
        w = np.frombuffer(data, dtype="<u8", count=n, offset=off)
        off += 8 * n
        return w

    hdr = (u64(), u64(), u64())
    rice = vec()
    ef = [u64() for _ in range(6)]
    vecs = [vec() for _ in range(4)]
    assert off == len(data)
    return hdr, rice, ef, vecs


def test_serialized_layout():
    keys = splitmix_keys(5, 10000)
    rs = OracleMPHF.build(keys, 8, 100)
    data = rs.serialize()
    (leaf, b, n), rice, ef, vecs = _parse(data)
    assert (leaf, b, n) == (8, 100, 10000)
    bits = int(lib.orc_rice_bits(rs.h))
    assert rice.shape[0] == ((bits + 7) // 8 + 14) // 8
    # sentinel bit (appendFixed(1, 1)) is the last bit
    assert (int(rice[(bits - 1) // 64]) >> ((bits - 1) % 64)) & 1 == 1
    assert ef[0] == 100  # num_buckets
    # round trip through the loader
    rs2 = OracleMPHF.load(data, 8)
    assert np.array_equal(rs2.query(keys), rs.query(keys))


def test_seed_count_matches_memo_nodes():
    leaf, bucket = 10, 300
    keys = splitmix_keys(11, 9000)
    rs = OracleMPHF.build(keys, leaf, bucket)
    memo = golomb_memo(leaf)
    nb = (9000 + bucket - 1) // bucket
    sizes = np.bincount([lib.orc_remap128(int(f), nb) for f in keys[:, 0]], minlength=nb)
    expect = sum(int((memo[s] >> 16) & 0x7FF) for s in sizes if s > 1)
    assert rs.nseeds == expect
    assert len(rs.seeds()) == expect


def test_errors():
    k = splitmix_keys(3, 100)
    k[5] = k[6]
    with pytest.raises(OracleError, match="equal second"):
        OracleMPHF.build(k, 8, 10)
    with pytest.raises(OracleError, match="3000"):
        OracleMPHF.build(splitmix_keys(3, 3500), 8, 5000)


# --------------------------------------------------------------------------- Elias-Fano


@pytest.mark.parametrize("nb", [1, 2, 255, 256, 257, 1000, 16383, 16384, 16385, 40000])
def test_elias_fano_roundtrip(nb):
    rng = np.random.default_rng(nb)
    sizes = rng.integers(0, 200, size=nb)
    cum = np.zeros(nb + 1, dtype=np.uint64)
    cum[1:] = np.cumsum(sizes)
    cum[-1] = max(int(cum[-1]), 1)
    bits = rng.integers(0, 400, size=nb) + sizes * 2
    pos = np.zeros(nb + 1, dtype=np.uint64)
    pos[1:] = np.cumsum(bits)
    assert lib.orc_ef_roundtrip(ptr64(cum), ptr64(pos), nb) == 0


def test_cpu_plain_config_1m_l8_b100():
    """The reference's CPU-runnable workload: n=1,000,000 splitmix64(seed 42)
    keys, l=8, b=100, rotation on, 1 thread: build, serialize, bijection."""
    keys = splitmix_keys(42, 1_000_000)
    rs = OracleMPHF.build(keys, 8, 100, rotate=True, threads=1)
    data = rs.serialize()
    assert len(data) > 0
    assert rs.bijection_errors(keys) == 0
    bits_per_key = len(data) * 8 / 1_000_000
    assert 1.5 < bits_per_key < 2.5
