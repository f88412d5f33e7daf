"""Autotuner (rs_autotune) and tuning strings through the C ABI: the tuned context
builds the same bytes as the default one, the chosen tuning round-trips through
rs_ctx_get_tune / rs_ctx_set_tune, and RS_TUNE-style strings change the launch."""
import pytest

from oracle_lib import OracleMPHF, splitmix_keys

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gra():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd
    return gpurecsplit_amd


@pytest.mark.parametrize("leaf,bucket", [(8, 100), (16, 2000)])
def test_autotune_keeps_bytes(gra, leaf, bucket):
    ctx = gra.Context(0)
    rep = ctx.autotune(leaf, bucket)
    print(rep["tune"], rep["search_ms_before"], rep["search_ms_after"], flush=True)
    assert rep["leaf"] == leaf and rep["sample_keys"] >= 2048 and len(rep["trials"]) >= 9
    assert ctx.get_tune(leaf) == rep["tune"]
    keys = splitmix_keys(leaf + 1000, 30000)
    got = gra.GPURecSplit(keys, bucket, leaf_size=leaf, ctx=ctx).serialize()
    assert got == OracleMPHF.build(keys, leaf, bucket, threads=8).serialize()
    other = gra.Context(0)
    other.set_tune(rep["tune"], leaf)
    assert other.get_tune(leaf) == rep["tune"]
    other.set_tune("u=1,g1=1,g2=1,bpc=4", leaf)
    t = other.get_tune(leaf)
    assert "u2=1" in t and "bpc=4" in t
    assert gra.GPURecSplit(keys, bucket, leaf_size=leaf, ctx=other).serialize() == got


@pytest.mark.parametrize("leaf,bucket,n", [(8, 100, 20000), (12, 100, 20000), (16, 2000, 30000),
                                           (18, 50, 6000), (20, 900, 1800), (24, 24, 480)])
def test_kernels_without_byte_tables_keep_bytes(gra, leaf, bucket, n):
    """The increment-table and remap-multiply kernels that the byte tables replace
    by default (tuning keys bt, btw, btl off) build the oracle's bytes, and so do
    the byte-table kernels at two candidates per lane."""
    keys = splitmix_keys(leaf * 7 + n, n)
    exp = OracleMPHF.build(keys, leaf, bucket, threads=8).serialize()
    for tune in ("bt=0,btw=0,btl=0", "ubt=2,ubtw=2,u3=2"):
        ctx = gra.Context(0)
        ctx.set_tune(tune, leaf)
        print(f"leaf={leaf} bucket={bucket} n={n} tune={tune}", flush=True)
        assert gra.GPURecSplit(keys, bucket, leaf_size=leaf, ctx=ctx).serialize() == exp
        ctx.close()
