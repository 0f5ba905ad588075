"""Strong Lucas and Baillie-PSW batch parity: the HIP lane-group kernels
against the GMP oracle (or_strong_lucas, pinned to OEIS A217255, and
mpz_probab_prime_p), bit-exact."""
import math
import random

import numpy as np
import pytest

import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu

# OEIS A217255 below 10^5 (strong Lucas pseudoprimes, Selfridge's method A)
STRONG_LUCAS_PSP = [5459, 5777, 10877, 16109, 18971, 22499, 24569, 25199, 40309, 58519, 75077, 97439]
# strong pseudoprimes to bases 2, 3, 5 (, 7, ...): OEIS A014233 terms
SPSP_235 = [25326001, 3215031751, 2152302898747, 3474749660383, 341550071728321]


def check(engine, vals, kind):
    limbs = max(pgap.limbs_needed(v) for v in vals)
    arr = pgap.to_limbs(vals, limbs)
    got = engine.prp(arr, kind)
    want = ol.verdicts(arr, kind)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(hex(vals[i]), int(got[i]), int(want[i])) for i in bad[:5]]
    return got


def qr_everywhere(bits, rng, prime=False):
    """Odd N = 1 (mod 4 * every odd prime <= 67): Jacobi(N/a) = 1 for every
    odd a <= 69, so the parameter search runs past its perfect-square check."""
    M = 4
    for p in range(3, 68, 2):
        if all(p % q for q in range(3, p, 2)):
            M *= p
    while True:
        n = 1 + M * rng.getrandbits(bits - M.bit_length())
        if math.isqrt(n) ** 2 == n:
            continue
        if not prime or ol.probab_prime(n):
            return n


def test_small_range(engine):
    vals = list(range(0, 30000))
    got_l = check(engine, vals, pgap.LUCAS)
    got_b = check(engine, vals, pgap.BPSW)
    comp_l = [v for v in vals if got_l[v] and not ol.probab_prime(v)]
    assert comp_l == [v for v in STRONG_LUCAS_PSP if v < 30000]
    assert [v for v in vals if got_b[v]] == [v for v in vals if ol.probab_prime(v)]


def test_pseudoprimes_squares_and_special_forms(engine):
    rng = random.Random(3)
    vals = STRONG_LUCAS_PSP + SPSP_235
    vals += [1093 ** 2, 3511 ** 2, 1093 ** 2 * 3511 ** 2, 10007 ** 2, (2 ** 61 - 1) ** 2, (2 ** 127 - 1) ** 2]
    vals += [(rng.getrandbits(b) | 1) ** 2 for b in (100, 500, 1000, 1500, 2100)]
    vals += [(1 << p) - 1 for p in (11, 23, 29, 61, 89, 127, 521, 607, 1109, 1279, 2203, 2281)]
    vals += [(1 << (1 << n)) + 1 for n in (5, 6, 7, 8, 9, 10, 11)]
    vals += [(1 << b) - 1 for b in (200, 896, 1344)]  # N + 1 = 2^b: d = 1
    vals += [qr_everywhere(b, rng) for b in (400, 1000, 2000)]
    vals += [qr_everywhere(b, rng, prime=True) for b in (400, 1000)]
    got_l = check(engine, vals, pgap.LUCAS)
    got_b = check(engine, vals, pgap.BPSW)
    for v, b in zip(vals, got_b):
        assert b == (1 if ol.probab_prime(v) else 0), hex(v)
    assert got_l[-2:].all() and got_b[-2:].all()


@pytest.mark.parametrize("bits", [30, 64, 200, 700, 894, 1108, 1340, 1800, 2128, 2236, 3000, 3576, 4248, 4470])
def test_random_sizes(engine, bits):
    rng = random.Random(bits)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(60)]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    vals.append(x)
    got_l = check(engine, vals, pgap.LUCAS)
    got_b = check(engine, vals, pgap.BPSW)
    assert got_l[-1] == 1 and got_b[-1] == 1
    assert [int(g) for g in got_b] == [1 if ol.probab_prime(v) else 0 for v in vals]


@pytest.mark.parametrize("lanes", [8, 16, 64])
def test_lane_families_agree(lanes):
    rng = random.Random(lanes)
    vals = []
    for bits in (700, 2128):
        vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(120)]
        with pgap.Engine(0, lanes=lanes) as e:
            check(e, vals, pgap.LUCAS)
            assert e.last_launch().lanes_per_number == lanes


def test_config_families(engine):
    # candidates of the four families the scan serves, free of factors below
    # 1000 (about one in 40 to 250 of them prime)
    for P, D, mlo, mhi, seed in ((503, 210, 1, 2000, 1), (809, 1841070, 10_000_000, 12_000_000, 2),
                                 (1511, 2190, 1, 200_000, 3), (3001, 2190, 1, 100_000, 4)):
        K = ol.K(P, D)
        rng = random.Random(seed)
        vals = []
        while len(vals) < 48:
            m = rng.randrange(mlo, mhi)
            if math.gcd(m, D) != 1:
                continue
            n = m * K + rng.randrange(-15000, 15001)
            if n % 2 and all(n % p for p in range(3, 1000, 2)):
                vals.append(n)
        check(engine, vals, pgap.BPSW)
