"""The header-only C++ drop-in (include/gpurecsplit_amd/GPURecSplit.hpp), exercised
by tests/cpp/dropin_test the way the reference's own tests use GPURecSplit:
operator>> and operator() without a GPU, the constructor and operator<< with one."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OracleMPHF, splitmix_keys

BIN = Path(__file__).resolve().parent / "cpp" / "dropin_test"


def _run(*args):
    return subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("leaf,bucket,n", [(8, 100, 20000), (16, 2000, 4000), (12, 100, 1)])
def test_cpp_load_and_query(tmp_path, leaf, bucket, n):
    keys = splitmix_keys(31 + leaf, n)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    (tmp_path / "mphf.bin").write_bytes(OracleMPHF.build(keys, leaf, bucket).serialize())
    r = _run("load", tmp_path / "mphf.bin", tmp_path / "keys.bin", leaf)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "load ok" in r.stdout


def test_cpp_reads_one_structure_per_extraction(tmp_path):
    """operator>> consumes exactly one structure: two MPHFs written back to back,
    then trailing bytes, are read by two extractions and the bytes stay unread."""
    ka, kb = splitmix_keys(61, 3000), splitmix_keys(62, 777)
    (tmp_path / "a.bin").write_bytes(ka.astype("<u8").tobytes())
    (tmp_path / "b.bin").write_bytes(kb.astype("<u8").tobytes())
    blob = OracleMPHF.build(ka, 8, 100).serialize() + OracleMPHF.build(kb, 8, 100).serialize() + b"TRAILER"
    (tmp_path / "s.bin").write_bytes(blob)
    r = _run("load2", tmp_path / "s.bin", tmp_path / "a.bin", tmp_path / "b.bin", 8)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "load2 ok: n=3000,777 trailing=TRAILER" in r.stdout


def test_cpp_load_rejects_truncated_stream(tmp_path):
    keys = splitmix_keys(63, 2000)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    data = OracleMPHF.build(keys, 8, 100).serialize()
    (tmp_path / "mphf.bin").write_bytes(data[:-5])
    r = _run("load", tmp_path / "mphf.bin", tmp_path / "keys.bin", 8)
    assert r.returncode == 3 and "error 8" in r.stdout


def test_cpp_recsplit_rotate_load_and_query(tmp_path):
    """recsplit_rotate::RecSplit (64-bit remap) reads the oracle's bytes and queries
    a bijection."""
    keys = splitmix_keys(64, 20000)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    (tmp_path / "mphf.bin").write_bytes(OracleMPHF.build(keys, 16, 2000, rotate=3).serialize())
    r = _run("load64", tmp_path / "mphf.bin", tmp_path / "keys.bin", 16)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "load ok" in r.stdout


def test_cpp_load_rejects_leaf_mismatch(tmp_path):
    keys = splitmix_keys(5, 1000)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    (tmp_path / "mphf.bin").write_bytes(OracleMPHF.build(keys, 8, 100).serialize())
    r = _run("load", tmp_path / "mphf.bin", tmp_path / "keys.bin", 12)
    assert r.returncode == 3 and "error 8" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("leaf,bucket,n", [(8, 100, 50000), (12, 100, 20000)])
def test_cpp_recsplit_rotate_build_matches_oracle(tmp_path, leaf, bucket, n):
    keys = splitmix_keys(99 + leaf, n)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    r = _run("build64", tmp_path / "keys.bin", leaf, bucket, tmp_path / "out.bin")
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "out.bin").read_bytes() == OracleMPHF.build(keys, leaf, bucket, rotate=3, threads=8).serialize()


@pytest.mark.gpu
@pytest.mark.parametrize("leaf,bucket,n", [(8, 100, 100000), (16, 2000, 20000)])
def test_cpp_build_matches_oracle(tmp_path, leaf, bucket, n):
    keys = splitmix_keys(77 + leaf, n)
    (tmp_path / "keys.bin").write_bytes(keys.astype("<u8").tobytes())
    r = _run("build", tmp_path / "keys.bin", leaf, bucket, tmp_path / "out.bin")
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "out.bin").read_bytes() == OracleMPHF.build(keys, leaf, bucket, threads=8).serialize()
