"""String-key front end (SpookyHash V2 first_hash, the reference's
AbstractParallelRecSplit.hpp:250 / include/util/Hash128.h:32-36) without a GPU:
the library's SpookyHash reproduces the published self-test values of SpookyHash
V2 for inputs of 0..32 bytes (the short path, which every key under 192 bytes
takes); first_hash swaps the two halves as the reference does; string queries
on a structure built from the hashed keys are a bijection, through the C ABI and
the C++ drop-in."""
import ctypes as C
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpurecsplit_amd as gra
from oracle_lib import OracleMPHF

HERE = Path(__file__).resolve().parent
SELFTEST = json.loads((HERE / "golden" / "spooky_v2_selftest.json").read_text())


def spooky128(data: bytes, s1: int = 0, s2: int = 0):
    h1, h2 = C.c_uint64(s1), C.c_uint64(s2)
    gra.lib.rs_spooky128(data, len(data), C.byref(h1), C.byref(h2))
    return h1.value, h2.value


@pytest.mark.parametrize("length", sorted(int(k) for k in SELFTEST["hash32_by_length"]))
def test_spooky_v2_published_selftest_values(length):
    buf = bytes((i + SELFTEST["byte_of_index_plus"]) & 0xFF for i in range(length))
    h1, _ = spooky128(buf, SELFTEST["seed"], SELFTEST["seed"])
    assert f"{h1 & 0xFFFFFFFF:08x}" == SELFTEST["hash32_by_length"][str(length)]


def test_first_hash_swaps_halves_and_batch_agrees():
    keys = [b"", b"a", b"key-17", bytes(range(200)), "üñï".encode(), b"x" * 191, b"y" * 192, b"z" * 1000]
    got = gra.hash_strings(keys)
    for k, (first, second) in zip(keys, got):
        h0, h1 = spooky128(k)
        assert (int(first), int(second)) == (h1, h0)
        out = (C.c_uint64 * 2)()
        gra.lib.rs_first_hash(k, len(k), out)
        assert (out[0], out[1]) == (int(first), int(second))


def test_spooky_length_and_seed_sensitivity():
    """Every length 0..400 and both seeds change the hash (no collisions among them)."""
    buf = bytes(range(256)) * 2
    seen = set()
    for n in range(401):
        seen.add(spooky128(buf[:n]))
        seen.add(spooky128(buf[:n], 1, 0))
        seen.add(spooky128(buf[:n], 0, 1))
    assert len(seen) == 3 * 401


def test_string_queries_are_a_bijection():
    keys = [f"https://example.org/item/{i}" for i in range(20000)]
    hk = gra.hash_strings(keys)
    data = OracleMPHF.build(hk, 8, 100).serialize()
    rs = gra.GPURecSplit.deserialize(data, 8)
    vals = np.array([rs.query_string(k) for k in keys], dtype=np.uint64)
    assert np.array_equal(np.sort(vals), np.arange(len(keys), dtype=np.uint64))
    assert np.array_equal(vals, rs.query(hk))


def test_cpp_string_queries(tmp_path):
    keys = [f"key:{i}:{i * i}" for i in range(5000)]
    (tmp_path / "keys.txt").write_text("\n".join(keys) + "\n")
    (tmp_path / "mphf.bin").write_bytes(OracleMPHF.build(gra.hash_strings(keys), 12, 100).serialize())
    r = subprocess.run([str(HERE / "cpp" / "dropin_test"), "loadstr", str(tmp_path / "mphf.bin"), str(tmp_path / "keys.txt"), "12"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "loadstr ok: n=5000" in r.stdout
