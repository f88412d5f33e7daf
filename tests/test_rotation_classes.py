"""The rotation-leaf test of leaves of at most 12 keys compares least rotations: a
candidate's right mask has a rotation r with left | rotl(right, r) == 2^L - 1
(SIMDRecSplit.hpp:452-509, the reference's fit) exactly when the right mask and the
complement of the left mask have the same least rotation, for every pair of masks
that nl even and L - nl odd keys can set.  Checked here by brute force: exhaustively
for L <= 8 and on random masks of L keys' bits for L <= 12."""
import numpy as np


def rotl(x, r, L):
    full = (1 << L) - 1
    return x if r == 0 else ((x << r) | (x >> (L - r))) & full


def least_rotation(x, L):
    return min(rotl(x, r, L) for r in range(L))


def fits(left, right, L):
    full = (1 << L) - 1
    return any(left | rotl(right, r, L) == full for r in range(L))


def reachable(popcount, keys):
    """A mask that `keys` keys can set: at least one bit when there are keys, at
    most one bit per key."""
    return min(1, keys) <= popcount <= keys


def test_least_rotation_criterion_exhaustive_small_leaves():
    """Every pair of masks that nl even and L - nl odd keys can set."""
    for L in range(2, 9):
        full = (1 << L) - 1
        canon = [least_rotation(x, L) for x in range(1 << L)]
        pop = [bin(x).count("1") for x in range(1 << L)]
        for nl in range(L + 1):
            lefts = [x for x in range(1 << L) if reachable(pop[x], nl)]
            rights = [x for x in range(1 << L) if reachable(pop[x], L - nl)]
            for left in lefts:
                for right in rights:
                    assert fits(left, right, L) == (canon[right] == canon[~left & full]), (L, nl, left, right)


def test_least_rotation_criterion_random_masks():
    rng = np.random.default_rng(12)
    for L in range(9, 13):
        full = (1 << L) - 1
        canon = [least_rotation(x, L) for x in range(1 << L)]
        for _ in range(4000):
            nl = int(rng.integers(0, L + 1))
            # masks as the kernel builds them: nl keys set bits of the left mask,
            # L - nl keys bits of the right one (collisions allowed)
            left = int(np.bitwise_or.reduce(1 << rng.integers(0, L, nl), initial=0))
            right = int(np.bitwise_or.reduce(1 << rng.integers(0, L, L - nl), initial=0))
            assert fits(left, right, L) == (canon[right] == canon[~left & full]), (L, left, right)


def test_least_rotation_matches_the_kernels_iteration():
    """The kernel fills its table by rotating left one bit at a time and keeping the
    minimum; the result is the least rotation."""
    for L in (2, 5, 8, 12):
        full = (1 << L) - 1
        for x in range(1 << L):
            c = y = x
            for _ in range(1, L):
                y = ((y << 1) | (y >> (L - 1))) & full
                c = min(c, y)
            assert c == least_rotation(x, L)
