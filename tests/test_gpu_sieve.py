"""Device sieve and gap_sieve_hip: the unknown sets against the oracle sieve
and the reference's own unknown-file md5 sums (misc/tests.sh:62-94)."""
import hashlib
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIEVE = os.path.join(REPO, "prime-gap_amd", "gap_sieve_hip")


@pytest.mark.parametrize("P,D,SL,limit,ms", [
    (907, 2190, 11000, 100_000_000, [1, 37, 199]),
    (503, 210, 4000, 1_000_000, [1, 11, 1999, 10**9 + 1]),
    (1511, 2190, 40000, 10_000_000, [1, 7, 199999]),
    (809, 1841070, 15000, 1000, [11244911]),
    (13, 1, 500, 3, [1, 2, 3]),
])
def test_sieve_matches_oracle(engine, P, D, SL, limit, ms):
    fam = engine.family(P, D, SL)
    got = engine.sieve(fam, ms, limit, SL)
    sv = ol.Sieve(P, D, SL, limit)
    for m, (p, n) in zip(ms, got):
        wp, wn = sv(m)
        assert np.array_equal(p, wp) and np.array_equal(n, wn), m


def test_sieve_primes_above_2_32(engine):
    # Primes past 2^32 take the 64-bit path (2^64 mod p, 16-bit K mod p
    # steps).  Sieving to 4.5e9 must remove from the 4e9 lists exactly the
    # offsets a prime in (4e9, 4.5e9] divides; the oracle names that prime
    # for each offset and Python checks the division with big integers.
    P, D, SL, ms = 907, 2190, 11000, [1, 17, 10**9 + 7]
    lo, hi = 4_000_000_000, 4_500_000_000
    fam = engine.family(P, D, SL)
    low = engine.sieve(fam, ms, lo, SL)
    high = engine.sieve(fam, ms, hi, SL)
    sv = ol.Sieve(P, D, SL, hi, above=lo)
    K = ol.K(P, D)
    crossed = 0
    for m, (lp, ln), (hp, hn) in zip(ms, low, high):
        removed = sv.removed(m, lo)
        for x, p in removed:
            assert p > 2**32 or p > lo
            assert (m * K + x) % p == 0 and ol.probab_prime(p), (m, x, p)
        gone = {x for x, _ in removed}
        want_p = [d for d in lp if -int(d) not in gone]
        want_n = [d for d in ln if int(d) not in gone]
        assert list(hp) == want_p and list(hn) == want_n, m
        crossed += sum(1 for x, p in removed if p > 2**32)
    assert crossed > 0


@pytest.mark.parametrize("flag,md5", [(None, "15a5cbff7301262caf047028c05f0525"),
                                      ("--rle", "2ed7b0b9621dab602f204865360a3c56"),
                                      ("--bitcompressed", "6edd5f50c4f588890a168461a7240c47")])
def test_gap_sieve_hip_reference_files(tmp_path, flag, md5):
    # misc/tests.sh:62-94
    args = [SIEVE, "-p", "907", "-d", "2190", "--mstart", "1", "--minc", "200", "--max-prime", "100",
            "--sieve-length", "11000", "--save", "-qq"] + ([flag] if flag else [])
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    data = (tmp_path / "unknowns" / "907_2190_1_200_s11000_l100M.txt").read_bytes()
    assert hashlib.md5(data).hexdigest() == md5


def test_gap_sieve_hip_953(tmp_path):
    # misc/tests.sh:70-71: 953_1_1_10000_s8000_l20M.txt
    r = subprocess.run([SIEVE, "-u", "953_1_1_10000_s8000_l20M.txt", "-qq"], cwd=tmp_path, capture_output=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    data = (tmp_path / "unknowns" / "953_1_1_10000_s8000_l20M.txt").read_bytes()
    assert hashlib.md5(data).hexdigest() == "a8d85c965ddf2d9601e93c9920a9aa12"


def test_sieve_then_scan_roundtrip(engine):
    # sieve on the device, scan on the device, rows = oracle on the same lists
    P, D, SL = 907, 2190, 11000
    fam = engine.family(P, D, SL)
    ms = [m for m in range(1000, 1100) if math.gcd(m, D) == 1]
    lists = engine.sieve(fam, ms, 10_000_000, SL)
    for m, (p, n) in zip(ms, lists):
        assert engine.scan_submit(fam, m, p, n) == pgap.PGAP_OK
    rows = {}
    while engine.scan_in_flight():
        for r in engine.scan_poll(wait=True):
            rows[r.m] = r
    for m, (p, n) in list(zip(ms, lists))[::5]:
        want = ol.test_interval(P, D, m, SL, p, n, kind=0)
        assert (rows[m].prev_p, rows[m].next_p, rows[m].p_tests, rows[m].n_tests) == \
            (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"])
