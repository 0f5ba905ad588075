"""CPU model of the lane-group squaring's rows (row-wise CIOS over G lanes of
S radix-2^28 digits, as the device runs it): each lane adds its bottom
accumulator's bits above 28 to its own next position and hands the low 28
bits to the lane below, the group's top lane takes zero, and the frame
moves down one digit per row.

The accumulators are 64-bit on the device.  Every step adds non-negative
terms, so a run with every product at its largest value bounds every real
run: with lazy input digits (a lane's digits 0 and 1 below 2^29 and
2^28 + 2^10, the others below 2^28) and m at 2^28 - 1 on every row, the
largest accumulator stays below 2^64 at every width the kernels use and
at S = 51 digits per lane, the limit the squaring asserts (from the
simpler count S(2^58 + 2^56) + 2^37 < 2^64, which takes every product at
its lazy maximum); in this model it first passes 2^64 between S = 120 and
S = 127, so the assertion is conservative.  Real squarings of random odd N
with lazy digits give x^2 R^-1 mod N below 2N."""
import random

import pytest

MASK = (1 << 28) - 1


def rows(x, n, np_, G, S, force_max=False):
    """-> (frame value after the L rows, largest accumulator seen)."""
    L = G * S
    T = [0] * L
    top = 0
    for i in range(L):
        xi = x[i]
        for p in range(L):
            T[p] += xi * x[p]
        m = MASK if force_max else (T[0] * np_) & MASK
        for p in range(L):
            T[p] += m * n[p]
        top = max(top, max(T))
        lo = [0] * G
        for g in range(G):
            b = g * S
            T[b + 1] += T[b] >> 28
            lo[g] = T[b] & MASK
        top = max(top, max(T))
        if not force_max:
            assert lo[0] == 0  # m clears the bottom digit: the shift is exact
        new = [0] * L
        for p in range(L):
            if (p + 1) % S == 0:  # a lane's top position
                g = p // S
                new[p] = lo[g + 1] if g + 1 < G else 0
            else:
                new[p] = T[p + 1]
        T = new
    return sum(v << (28 * p) for p, v in enumerate(T)), top


def lazy_max_digits(G, S):
    d = []
    for g in range(G):
        for j in range(S):
            d.append((1 << 29) - 1 if j == 0 else ((1 << 28) + (1 << 10) - 1 if j == 1 else MASK))
    return d


@pytest.mark.parametrize("G,S", [(16, 10), (8, 19), (8, 20), (4, 38), (4, 51)])
def test_worst_case_accumulators_fit_64_bits(G, S):
    _, top = rows(lazy_max_digits(G, S), [MASK] * (G * S), 0, G, S, force_max=True)
    assert top < 1 << 64, (G, S, top.bit_length())


def test_the_model_sees_an_overflow():
    # the same worst case at 128 digits per lane does overflow: the check
    # above is not vacuous
    _, top = rows(lazy_max_digits(4, 128), [MASK] * (4 * 128), 0, 4, 128, force_max=True)
    assert top >= 1 << 64


def to_digits(v, L):
    return [(v >> (28 * i)) & MASK for i in range(L)]


@pytest.mark.parametrize("G,S", [(4, 38), (8, 19)])
def test_rows_compute_the_montgomery_square(G, S):
    rng = random.Random(G * 1000 + S)
    L = G * S
    R = 1 << (28 * L)
    for _ in range(3):
        N = rng.getrandbits(28 * L - 6) | (1 << (28 * L - 7)) | 1   # the variants' headroom
        np_ = (-pow(N, -1, 1 << 28)) % (1 << 28)
        x = rng.randrange(0, 4 * N)
        # lazy digits of the same value: one unit of each lane's digit 1
        # moved into its digit 0 (digit 0 then above 28 bits, below 2^29)
        d = to_digits(x, L)
        for g in range(G):
            b = g * S
            if d[b + 1]:
                d[b + 1] -= 1
                d[b] += 1 << 28
        assert sum(v << (28 * i) for i, v in enumerate(d)) == x and max(d) < 1 << 29
        v, top = rows(d, to_digits(N, L), np_, G, S)
        assert top < 1 << 64
        assert v % N == x * x * pow(R, -1, N) % N
        assert v < 2 * N
