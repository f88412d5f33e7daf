"""Allocation-free build entry point (rs_build_device_noalloc), called through the
C ABI: its serialized output equals the serialized result of the existing entry
point rs_build_device on the same keys, on two calls in a row (the second with a
smaller input, so stale scratch contents from the first must not leak), and the
device's free memory does not change across its calls."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpurecsplit_amd as gra
    from gpurecsplit_amd.keys import splitmix64_keys_device
    return torch, gra, splitmix64_keys_device


def _noalloc(gra, ctx, d, bucket, leaf, cap):
    buf = C.create_string_buffer(cap)
    ln = C.c_size_t(0)
    rc = gra.lib.rs_build_device_noalloc(ctx.h, C.c_void_p(d.data_ptr()), d.shape[0], leaf, bucket, 1, buf, cap,
                                         C.byref(ln))
    return rc, buf.raw[: ln.value], ln.value


def test_noalloc_equals_build_device_twice(env):
    torch, gra, gen = env
    ctx = gra.Context(0)
    cases = [(300_000, 8, 100, 1), (40_000, 16, 2000, 2)]  # second call smaller
    inputs = [(gen(seed, n), leaf, bucket) for n, leaf, bucket, seed in cases]
    torch.cuda.synchronize()
    expected = []
    for d, leaf, bucket in inputs:
        print(f"rs_build_device n={d.shape[0]} l={leaf} b={bucket}", flush=True)
        expected.append(gra.GPURecSplit(d, bucket, leaf_size=leaf, ctx=ctx).serialize())
    # warm-up call: the runtime may set up its own transfer staging on first use
    rc, got, _ = _noalloc(gra, ctx, inputs[1][0], inputs[1][2], inputs[1][1], len(expected[1]))
    assert rc == 0 and got == expected[1]
    free0 = torch.cuda.mem_get_info()[0]
    for (d, leaf, bucket), exp in zip(inputs, expected):
        print(f"rs_build_device_noalloc n={d.shape[0]} l={leaf} b={bucket}", flush=True)
        rc, got, ln = _noalloc(gra, ctx, d, bucket, leaf, len(exp) + 64)
        assert rc == 0, gra.lib.rs_last_error()
        assert ln == len(exp) and got == exp
    assert torch.cuda.mem_get_info()[0] == free0
    assert gra.build_noalloc(inputs[0][0], 100, 8, ctx=ctx) == expected[0]
    sys.stdout.flush()


def test_noalloc_rejects_oversized_input_and_small_buffer(env):
    torch, gra, gen = env
    ctx = gra.Context(0)
    d = gen(3, 1_000_000)
    rc, _, _ = _noalloc(gra, ctx, d, 1, 8, 1 << 20)  # 1M buckets of 1 key: beyond the 64 MiB scratch
    assert rc == 1 and b"scratch" in gra.lib.rs_last_error()
    small = d[:5000]
    exp = gra.GPURecSplit(small, 100, leaf_size=8, ctx=ctx).serialize()
    rc, _, need = _noalloc(gra, ctx, small, 100, 8, 16)
    assert rc == 1 and need == len(exp)
    rc, got, _ = _noalloc(gra, ctx, small, 100, 8, need)
    assert rc == 0 and got == exp
