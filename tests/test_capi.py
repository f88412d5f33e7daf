"""CPU checks of the product library: it loads, exports the C ABI declared in
include/gpurecsplit_amd.h, and its host-side code (memo, deserialization, host
query, serialization) agrees with the oracle without touching a GPU."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

import gpurecsplit_amd as gra
from oracle_lib import OracleMPHF, golomb_memo as oracle_memo, splitmix_keys

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "gpurecsplit_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    raw = C.CDLL(str(gra.LIB_PATH))
    for s in syms:
        assert hasattr(raw, s), s


@pytest.mark.parametrize("leaf", range(2, 25))
def test_product_memo_matches_reference_fixture(leaf):
    fx = json.loads((ROOT / "tests" / "golden" / "golomb_memo.json").read_text())[str(leaf)]
    memo = gra.golomb_memo(leaf)
    for m, v in fx["samples"].items():
        assert int(memo[int(m)]) == v
    assert np.array_equal(memo, oracle_memo(leaf))


@pytest.mark.parametrize("leaf,bucket,n,rotate", [(8, 100, 20000, True), (5, 7, 3000, False), (16, 2000, 3000, True),
                                                  (12, 60, 1, True), (3, 2, 101, True), (24, 10, 30, True)])
def test_host_query_and_serialization_on_oracle_bytes(leaf, bucket, n, rotate):
    keys = splitmix_keys(1000 + leaf, n)
    ref = OracleMPHF.build(keys, leaf, bucket, rotate=rotate)
    data = ref.serialize()
    rs = gra.GPURecSplit.deserialize(data, leaf, rotate=rotate)
    assert rs.serialize() == data
    got = rs.query(keys)
    assert np.array_equal(got, ref.query(keys))
    assert np.array_equal(np.sort(got), np.arange(n, dtype=np.uint64))
    assert rs(tuple(int(v) for v in keys[0])) == int(got[0])
    assert rs.num_keys == n


def test_deserialize_rejects_bad_streams():
    keys = splitmix_keys(3, 500)
    data = OracleMPHF.build(keys, 8, 50).serialize()
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit.deserialize(data, 9)
    assert e.value.code == 8
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit.deserialize(data[: len(data) // 2], 8)
    assert e.value.code == 8


def _vec_offsets(data: bytes):
    """Byte offsets of the length words of the five vectors of a serialized MPHF."""
    import struct
    offs, p = [], 24
    for i in range(5):
        offs.append(p)
        n = struct.unpack_from("<Q", data, p)[0]
        p += 8 + 8 * n
        if i == 0:
            p += 48  # the six Elias-Fano scalars follow the descriptor vector
    assert p == len(data)
    return offs


def test_deserialize_rejects_inconsistent_headers():
    import struct
    keys = splitmix_keys(21, 5000)
    data = OracleMPHF.build(keys, 8, 100).serialize()
    rice_end = _vec_offsets(data)[1] - 48
    cases = {
        "key count": struct.pack("<Q", 7000),             # bucket count no longer matches the EF
        "bucket size": struct.pack("<Q", 10),
    }
    for name, patch in cases.items():
        off = 16 if name == "key count" else 8
        bad = data[:off] + patch + data[off + 8:]
        with pytest.raises(gra.RecSplitError) as e:
            gra.GPURecSplit.deserialize(bad, 8)
        assert e.value.code == 8, name
    # EF bucket count changed
    bad = data[:rice_end] + struct.pack("<Q", 49) + data[rice_end + 8:]
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit.deserialize(bad, 8)
    assert e.value.code == 8
    # one word cut from each EF vector (length word and data shortened together)
    for vi in range(1, 5):
        o = _vec_offsets(data)[vi]
        n = struct.unpack_from("<Q", data, o)[0]
        bad = data[:o] + struct.pack("<Q", n - 1) + data[o + 8:o + 8 + 8 * (n - 1)] + data[o + 8 + 8 * n:]
        with pytest.raises(gra.RecSplitError) as e:
            gra.GPURecSplit.deserialize(bad, 8)
        assert e.value.code == 8, f"vector {vi}"
    # empty descriptor vector
    o = _vec_offsets(data)[0]
    n = struct.unpack_from("<Q", data, o)[0]
    bad = data[:o] + struct.pack("<Q", 0) + data[o + 8 + 8 * n:]
    with pytest.raises(gra.RecSplitError) as e:
        gra.GPURecSplit.deserialize(bad, 8)
    assert e.value.code == 8


def test_host_query_of_corrupted_contents_stays_in_bounds():
    """Random words in the descriptor and Elias-Fano arrays: either the stream is
    rejected, or every host query returns (in bounds of its buffers) some value."""
    keys = splitmix_keys(22, 3000)
    data = bytearray(OracleMPHF.build(keys, 8, 100).serialize())
    rng = np.random.default_rng(0)
    offs = _vec_offsets(bytes(data))
    import struct
    for trial in range(40):
        bad = bytearray(data)
        vi = trial % 5
        n = struct.unpack_from("<Q", bad, offs[vi])[0]
        for _ in range(4):
            w = int(rng.integers(0, max(1, n)))
            struct.pack_into("<Q", bad, offs[vi] + 8 + 8 * w, int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2)))
        try:
            rs = gra.GPURecSplit.deserialize(bytes(bad), 8)
        except gra.RecSplitError as e:
            assert e.code == 8
            continue
        rs.query(keys)  # must not crash


def test_bits_count_matches_reference_formula():
    """getBits() = EF bits (DoubleEF.hpp:435-437) + descriptor words * 64."""
    keys = splitmix_keys(4, 10000)
    data = OracleMPHF.build(keys, 8, 100).serialize()
    rs = gra.GPURecSplit.deserialize(data, 8)
    bits = rs.get_bits()
    assert 1.4 * 10000 < bits < 3 * 10000


def test_time_limit_rejects_null_context():
    assert gra.lib.rs_ctx_set_time_limit(None, 1.0) == 1
    assert gra.lib.rs_ctx_set_time_limit(None, -1.0) == 1


@pytest.mark.parametrize("leaf,bucket,n,flags", [(8, 100, 20000, 3), (16, 2000, 4000, 3), (5, 7, 3000, 2), (2, 7, 500, 3)])
def test_host_query_remap64_on_oracle_bytes(leaf, bucket, n, flags):
    """The RecSplitRotate variant (64-bit remap): a structure serialized by the
    oracle loads with remap64 and answers every query as the oracle does."""
    keys = splitmix_keys(2000 + leaf, n)
    ref = OracleMPHF.build(keys, leaf, bucket, rotate=flags)
    data = ref.serialize()
    rs = gra.GPURecSplit.deserialize(data, leaf, rotate=bool(flags & 1), remap64=True)
    assert rs.serialize() == data
    got = rs.query(keys)
    assert np.array_equal(got, ref.query(keys))
    assert np.array_equal(np.sort(got), np.arange(n, dtype=np.uint64))
    # the 16-bit reading of the same bytes is a different function
    assert not np.array_equal(gra.GPURecSplit.deserialize(data, leaf, rotate=bool(flags & 1)).query(keys), got)


def test_tune_strings_without_gpu():
    """Tuning strings are C ABI data: rs_ctx_set_tune / rs_ctx_get_tune need a
    context (a GPU); without one, context creation fails cleanly."""
    h = C.c_void_p()
    rc = gra.lib.rs_ctx_create(0, C.byref(h))
    if rc == 0:
        gra.lib.rs_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == 2 and b"device" in gra.lib.rs_last_error()
