"""Pins the GMP oracle against facts it does not compute itself.

- verdicts: Python's own big-integer pow() (an independent bignum library),
  trial-division primality on small N, published pseudoprime lists (OEIS
  A001567 prefix, A001262 below 10^5), Mersenne and Fermat numbers;
- sieve + line formats: md5 sums the reference's integration test pins
  (misc/tests.sh:66-94) and unknown counts printed in README.md:245-262;
- scan rows: README.md:265-288 (prev/next/merit per m) and the prp-test
  totals of README.md:241-262;
- K: a Python product of primes; the record gap of README.md:40-46.
"""
import hashlib
import math
import random

import numpy as np
import pytest

import oracle_lib as ol
import unknown_format as uf


def is_prime_td(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    f = 3
    while f * f <= n:
        if n % f == 0:
            return False
        f += 2
    return True


def primes_upto(n):
    return [p for p in range(2, n + 1) if is_prime_td(p)]


# ---------------------------------------------------------------- verdicts

def test_fermat2_small_brute_force():
    psp = []
    for n in range(0, 20000):
        v = ol.fermat2(n)
        expect = 1 if (n % 2 == 1 and n > 1 and pow(2, n - 1, n) == 1) else 0
        assert v == expect, n
        if v and not is_prime_td(n):
            psp.append(n)
        if n > 2 and n % 2 == 1 and is_prime_td(n):
            assert v == 1, n
    # OEIS A001567 (base-2 Fermat pseudoprimes) starts:
    assert psp[:8] == [341, 561, 645, 1105, 1387, 1729, 1905, 2047]


def test_sprp_small_published_lists():
    spsp2 = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581, 85489,
             88357, 90751]  # OEIS A001262 below 10^5
    got = [n for n in range(3, 100000, 2) if ol.sprp(n, 2) and not is_prime_td(n)]
    assert got == spsp2
    for n in range(3, 3000, 2):
        assert ol.sprp(n, 2) == (1 if is_prime_td(n) or n in spsp2 else 0), n
    # strong pseudoprime to bases 2, 3, 5 (and 25326001 is the least such)
    assert all(ol.sprp(25326001, b) for b in (2, 3, 5)) and not is_prime_td(25326001)
    assert all(ol.sprp(3215031751, b) for b in (2, 3, 5, 7))
    assert ol.sprp(3215031751, 11) == 0
    assert ol.sprp(1, 2) == 0 and ol.sprp(4, 3) == 0 and ol.sprp(3, 3) == 0 and ol.sprp(5, 3) == 1


def test_probab_prime_small():
    for n in range(0, 30000):
        assert (ol.probab_prime(n) > 0) == is_prime_td(n), n


# OEIS A217255: strong Lucas pseudoprimes with Selfridge's method A, below 10^5
STRONG_LUCAS_PSP = [5459, 5777, 10877, 16109, 18971, 22499, 24569, 25199, 40309, 58519, 75077, 97439]


def jacobi_py(a: int, n: int) -> int:
    """Jacobi symbol (a/n), n odd > 0 (textbook reciprocity loop)."""
    a %= n
    t = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                t = -t
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            t = -t
        a %= n
    return t if n == 1 else 0


def strong_lucas_linear(n: int) -> int:
    """Strong Lucas test by direct linear recurrences U_k+1 = P U_k - Q U_k-1,
    V likewise (no doubling formulas): independent of the oracle's chain."""
    if n < 2 or (n % 2 == 0 and n != 2):
        return 0
    if n == 2:
        return 1
    if math.isqrt(n) ** 2 == n:
        return 0
    a = 5
    while True:
        D = -a if a % 4 == 3 else a
        j = jacobi_py(D, n)
        if j == -1:
            break
        if j == 0 and not (n <= a and a % n == 0):
            return 0
        a += 2
    P, Q = 1, (1 - D) // 4
    d, s = n + 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    U = [0, 1]
    V = [2, P % n]
    for _ in range(2, d * (1 << (s - 1)) + 1):
        U.append((P * U[-1] - Q * U[-2]) % n)
        V.append((P * V[-1] - Q * V[-2]) % n)
    if U[d] == 0 or V[d] == 0:
        return 1
    return 1 if any(V[d << r] == 0 for r in range(1, s)) else 0


def test_strong_lucas_published_list_and_primes():
    got = [n for n in range(1, 100000) if ol.strong_lucas(n) and not is_prime_td(n)]
    assert got == STRONG_LUCAS_PSP
    assert all(ol.strong_lucas(p) for p in primes_upto(100000))
    # squares (no Selfridge D exists) and even numbers fail; so does 1
    for n in (1, 4, 9, 25, 49, 1093 ** 2, 3511 ** 2, 10007 ** 2, (2 ** 61 - 1) ** 2):
        assert ol.strong_lucas(n) == 0, n


def test_strong_lucas_linear_recurrence():
    for n in range(1, 2500):
        assert ol.strong_lucas(n) == strong_lucas_linear(n), n
    for n in STRONG_LUCAS_PSP[:4]:
        assert strong_lucas_linear(n) == 1


def test_bpsw_equals_probab_prime():
    for n in range(0, 200000):
        assert ol.bpsw(n) == (1 if ol.probab_prime(n) else 0), n
    # strong base-2 pseudoprimes that also pass bases 3, 5, 7 (OEIS A014233
    # terms) and the strong Lucas pseudoprimes: composite, rejected by BPSW
    for n in [25326001, 3215031751, 2152302898747, 3474749660383, 341550071728321] + STRONG_LUCAS_PSP:
        assert ol.bpsw(n) == 0 and ol.probab_prime(n) == 0, n
    assert ol.sprp(341550071728321, 2) and ol.sprp(341550071728321, 13)
    rng = random.Random(77)
    for bits in (700, 1108, 2128, 4248):
        vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(40)]
        x = vals[0]
        while not ol.probab_prime(x):
            x += 2
        vals.append(x)
        for v in vals:
            assert ol.bpsw(v) == (1 if ol.probab_prime(v) else 0)
            assert ol.strong_lucas(v) == (1 if ol.probab_prime(v) else 0)
    # composite Mersenne numbers are strong base-2 pseudoprimes; Lucas rejects them
    for p in (11, 23, 1109, 2131):
        assert ol.sprp(2 ** p - 1, 2) and ol.bpsw(2 ** p - 1) == 0
    for p in (521, 607, 1279, 2203):
        assert ol.bpsw(2 ** p - 1) == 1


@pytest.mark.parametrize("bits", [33, 64, 65, 127, 700, 1108, 1500, 2130])
def test_fermat2_vs_python_pow(bits):
    rng = random.Random(1000 + bits)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(60)]
    # make some of them prime so both verdicts occur (candidates located with
    # the oracle, verdicts checked with Python's pow)
    primes = []
    for v in vals[:4]:
        x = v
        while not ol.probab_prime(x):
            x += 2
        assert pow(3, x - 1, x) == 1
        primes.append(x)
    for x in vals + primes:
        assert ol.fermat2(x) == (1 if pow(2, x - 1, x) == 1 else 0)
    assert all(ol.fermat2(x) for x in primes)


def test_mersenne_and_fermat_numbers():
    for p in (521, 607, 1279, 2203, 2281):
        m = (1 << p) - 1
        assert ol.fermat2(m) == 1 and ol.sprp(m, 2) == 1 and ol.sprp(m, 3) == 1 and ol.probab_prime(m) > 0
    for p in (11, 23, 29, 37, 41, 43, 47, 53, 59, 67, 1109):
        m = (1 << p) - 1   # composite Mersenne numbers with prime exponent: strong psp base 2
        assert ol.probab_prime(m) == 0
        assert ol.fermat2(m) == 1 and ol.sprp(m, 2) == 1
        assert ol.sprp(m, 3) == 0
    for n in (5, 10, 11, 12):
        f = (1 << (1 << n)) + 1  # F_n: 2^(F_n - 1) == 1 mod F_n for every n
        assert ol.fermat2(f) == 1 and ol.sprp(f, 2) == 1
        assert ol.probab_prime(f) == 0
        assert ol.fermat2(f + 2) == (1 if pow(2, f + 1, f + 2) == 1 else 0)


def test_batched_verdicts_match_single():
    rng = random.Random(7)
    vals = [rng.getrandbits(300) | 1 for _ in range(200)] + [(1 << 521) - 1, 341, 561, 2047, 0, 1, 2, 3]
    limbs = max((v.bit_length() + 31) // 32 for v in vals)
    arr = np.zeros((len(vals), limbs), dtype=np.uint32)
    for i, v in enumerate(vals):
        arr[i] = np.frombuffer(v.to_bytes(4 * limbs, "little"), dtype="<u4")
    f = ol.verdicts(arr, 0, threads=4)
    s = ol.verdicts(arr, 1, bases=(2, 3), threads=3)
    for i, v in enumerate(vals):
        assert f[i] == ol.fermat2(v)
        assert s[i] == (ol.sprp(v, 2) and ol.sprp(v, 3))


# ---------------------------------------------------------------- K, gaps

def test_K_and_K_log():
    for P, D in ((503, 210), (809, 1841070), (907, 2190), (1511, 2190), (3001, 2190), (13, 1)):
        k = math.prod(primes_upto(P)) // D
        assert ol.K(P, D) == k
        assert abs(ol.K_log(P, D) - math.log(k)) < 1e-9 * math.log(k)
    with pytest.raises(ValueError):
        ol.K(503, 11 * 1009)


def test_prev_next_prime_small_family():
    K = math.prod(primes_upto(11))  # 2310, D = 1
    for m in range(1, 60):
        c = m * K
        for start in (0, 5, 50):
            nx = c + start + 1
            while not is_prime_td(nx):
                nx += 1
            assert ol.next_prime_dist(11, 1, m, start) == nx - c
            pv = c - start - 1
            while not is_prime_td(pv):
                pv -= 1
            assert ol.prev_prime_dist(11, 1, m, start) == c - pv


def test_readme_record_gap():
    # README.md:40-46: 23142 30.1485  11244911 * 809#/1841070 -5788 to +17354
    m, P, D = 11244911, 809, 1841070
    assert ol.prev_prime_dist(P, D, m, 0) == 5788
    assert ol.next_prime_dist(P, D, m, 0) == 17354
    merit = np.float32((5788 + 17354) / (ol.K_log(P, D) + math.log(m)))
    assert f"{merit:.4f}" == "30.1485"


# ---------------------------------------------------------------- sieve, formats, scan

def test_sieve_reproduces_reference_unknown_file(unknowns_907):
    s = ol.Sieve(907, 2190, 11000, 100_000_000)
    lines = []
    for m in range(1, 201):
        if math.gcd(m, 2190) != 1:
            continue
        p, n = s(m)
        lines.append(uf.plain_line(m, p, n))
    data = b"".join(lines)
    # misc/tests.sh:66-67
    assert hashlib.md5(data).hexdigest() == "15a5cbff7301262caf047028c05f0525"
    assert data == unknowns_907


def test_formats_reproduce_reference_md5(unknowns_907):
    K = ol.K(907, 2190)
    cop = uf.coprime_x(K, 11000)
    rle, bc = [], []
    for ln in unknowns_907.splitlines(keepends=True):
        m, p, n = uf.parse_line(ln)
        rle.append(uf.rle_line(m, p, n))
        bc.append(uf.bitcompressed_line(m, p, n, K, 2190, 11000, cop))
        assert uf.parse_line(rle[-1]) == (m, p, n)
        assert uf.parse_line(bc[-1], K, 2190, 11000, cop) == (m, p, n)
    assert hashlib.md5(b"".join(rle)).hexdigest() == "2ed7b0b9621dab602f204865360a3c56"  # misc/tests.sh:80,85
    assert hashlib.md5(b"".join(bc)).hexdigest() == "6edd5f50c4f588890a168461a7240c47"   # misc/tests.sh:89,94


README_ROWS = {  # README.md:265-288 (gap_test.py output, prev/next as GMP confirms)
    1: (7300, 450, "9.0185"), 17: (480, 6982, "8.6548"), 19: (216, 6722, "8.0460"),
    53: (4436, 4084, "9.8689"), 143: (6582, 1800, "9.6979"), 163: (7246, 432, "8.8820"),
    37: (250, 1296, None), 113: (972, 40, None), 199: (30, 192, None),
}
README_UNKNOWNS = {1: (218, 212), 37: (214, 234), 113: (207, 215), 199: (216, 205)}
README_TOTALS = {1: (430, 144), 10: (4262, 643), 30: (12676, 1550), 53: (22377, 2590)}  # after k m's


def test_scan_rows_match_readme(unknowns_907):
    K_log = ol.K_log(907, 2190)
    tests = unknowns = 0
    for k, ln in enumerate(unknowns_907.splitlines(keepends=True), start=1):
        m, p, n = uf.parse_line(ln)
        r = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        tests += r["tests"]
        unknowns += len(p) + len(n)
        if m in README_UNKNOWNS:
            assert (len(p), len(n)) == README_UNKNOWNS[m]
        if m in README_ROWS:
            prev, nxt, merit = README_ROWS[m]
            assert (r["prev_p"], r["next_p"]) == (prev, nxt), m
            if merit:
                mf = np.float32((prev + nxt) / (K_log + math.log(m)))
                assert f"{mf:.4f}" == merit
        if k in README_TOTALS:
            assert (unknowns, tests) == README_TOTALS[k], k
        # the Fermat-2 scan gives the same row on these inputs
        r2 = ol.test_interval(907, 2190, m, 11000, p, n, kind=1)
        assert (r2["prev_p"], r2["next_p"], r2["tests"]) == (r["prev_p"], r["next_p"], r["tests"])


def test_scan_fallback_beyond_sieve_length():
    # README.md:40-46 record: next prime is 17354 > SL = 15000, so an empty
    # next list forces the nextprime fallback (gap_test_common.cpp:170-178).
    m, P, D = 11244911, 809, 1841070
    r = ol.test_interval(P, D, m, 15000, [5788], [], kind=0)
    assert (r["prev_p"], r["next_p"], r["next_fallback"], r["prev_fallback"]) == (5788, 17354, 1, 0)
    r = ol.test_interval(P, D, m, 5000, [], [], kind=0)
    assert r["prev_p"] == 5788 and r["prev_fallback"] == 1 and r["tests"] == 0
