"""Synthetic key generators of the package: host (numpy) and device (torch)
SplitMix64 streams equal the oracle's generator, whose first outputs are pinned
by the SplitMix64 known answers in test_oracle.py."""
import numpy as np
import pytest

from gpurecsplit_amd.keys import splitmix64_keys, splitmix64_keys_device
from oracle_lib import splitmix_keys


@pytest.mark.parametrize("seed,n", [(0, 5), (1, 1000), (42, 4097), (2**63 + 5, 300)])
def test_host_generator_matches_oracle(seed, n):
    assert np.array_equal(splitmix64_keys(seed, n), splitmix_keys(seed, n))


def test_generator_offsets_and_chunks_agree():
    full = splitmix64_keys(3, 5000)
    assert np.array_equal(splitmix64_keys(3, 700, start=1234), full[1234:1934])
    torch = pytest.importorskip("torch")
    dev = splitmix64_keys_device(3, 5000, device="cpu", chunk=999).numpy().view(np.uint64)
    assert np.array_equal(dev, full)
    part = splitmix64_keys_device(3, 700, device="cpu", start=1234).numpy().view(np.uint64)
    assert np.array_equal(part, full[1234:1934])
