"""String keys on the GPU: the device SpookyHash V2 kernel equals the host one for
every input length 0..700 (short and long paths), and MPHFs built from string
keys (rs_build_strings, the C++ vector<string> / ifstream constructors) are
byte-identical to the oracle's build of the same keys' first_hash values."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OracleMPHF

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd as gra
    return torch, gra


def test_device_hash_equals_host_hash(env):
    torch, gra = env
    rng = np.random.default_rng(7)
    keys = [rng.integers(0, 256, size=n, dtype=np.uint8).tobytes() for n in range(701)]
    blob = b"".join(keys)
    off = np.zeros(len(keys) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(k) for k in keys])
    d_bytes = torch.from_numpy(np.frombuffer(blob, dtype=np.uint8).copy()).cuda()
    d_off = torch.from_numpy(off.view(np.int64)).cuda()
    d_out = torch.empty((len(keys), 2), dtype=torch.int64, device="cuda")
    ctx = gra.Context.default()
    torch.cuda.synchronize()
    assert gra.lib.rs_hash_strings_device(ctx.h, C.c_void_p(d_bytes.data_ptr()), C.c_void_p(d_off.data_ptr()),
                                          len(keys), C.c_void_p(d_out.data_ptr())) == 0
    assert np.array_equal(d_out.cpu().numpy().view(np.uint64), gra.hash_strings(keys))


@pytest.mark.parametrize("leaf,bucket,n", [(8, 100, 50_000), (16, 2000, 20_000)])
def test_string_build_matches_oracle(env, leaf, bucket, n):
    torch, gra = env
    keys = [f"https://example.org/{leaf}/item/{i}?q={i * 7919 % 1000}" for i in range(n)]
    rs = gra.GPURecSplit.from_strings(keys, bucket, leaf_size=leaf)
    exp = OracleMPHF.build(gra.hash_strings(keys), leaf, bucket, threads=8).serialize()
    assert rs.serialize() == exp
    vals = np.array([rs.query_string(k) for k in keys[:2000]], dtype=np.uint64)
    assert len(set(vals.tolist())) == 2000 and vals.max() < n


def test_cpp_ifstream_constructor(env, tmp_path):
    torch, gra = env
    keys = [f"line {i} of the key file" for i in range(30_000)]
    (tmp_path / "keys.txt").write_text("\n".join(keys) + "\n")
    r = subprocess.run([str(HERE / "cpp" / "dropin_test"), "buildstr", str(tmp_path / "keys.txt"), "12", "100",
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "buildstr ok: n=30000" in r.stdout
    exp = OracleMPHF.build(gra.hash_strings(keys), 12, 100, threads=8).serialize()
    assert (tmp_path / "out.bin").read_bytes() == exp
