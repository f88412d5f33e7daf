"""Synthetic key sets: uniform 128-bit key hashes from SplitMix64.

Key i of SplitMix64(seed) is (first, second) = (next(), next()): the generator
state advances by the golden-ratio increment and each output is the Stafford-13
finalizer of the state (the same finalizer as `remix`,
AbstractParallelRecSplit.hpp:265-269).  The host and device generators produce
the same keys.
"""
from __future__ import annotations

import numpy as np

GAMMA = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB


def splitmix64_keys(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Keys start .. start+n-1 of SplitMix64(seed) as an (n, 2) uint64 array."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(GAMMA) * np.arange(2 * start + 1, 2 * (start + n) + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
        z = z ^ (z >> np.uint64(31))
    return z.reshape(n, 2)


def _s64(x: int) -> int:
    """The int64 with the same 64 bits as the unsigned x."""
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >= 1 << 63 else x


def splitmix64_keys_device(seed: int, n: int, device="cuda:0", start: int = 0, chunk: int = 1 << 26):
    """Same keys generated on the GPU as an (n, 2) int64 tensor (bit patterns of
    the uint64 keys).  torch's int64 arithmetic wraps modulo 2^64; right shifts
    are made logical by masking the sign-extended bits."""
    import torch

    out = torch.empty((n, 2), dtype=torch.int64, device=device)
    flat = out.view(-1)
    m1, m2, g = _s64(M1), _s64(M2), _s64(GAMMA)
    for lo in range(0, 2 * n, chunk):
        hi = min(2 * n, lo + chunk)
        z = torch.arange(2 * start + lo + 1, 2 * start + hi + 1, dtype=torch.int64, device=device)
        z.mul_(g).add_(_s64(seed))
        z = (z ^ ((z >> 30) & ((1 << 34) - 1))) * m1
        z = (z ^ ((z >> 27) & ((1 << 37) - 1))) * m2
        flat[lo:hi] = z ^ ((z >> 31) & ((1 << 33) - 1))
    return out
