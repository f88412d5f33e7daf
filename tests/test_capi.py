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
