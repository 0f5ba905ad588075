"""Device-side candidate construction and trial-division screening."""
import math
import random

import numpy as np
import pytest

import hipmem
import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu


def test_build_candidates_matches_python(engine):
    fam = engine.family(809, 1841070, 0)
    K = ol.K(809, 1841070)
    rng = random.Random(3)
    m = np.array([rng.randrange(1, 1 << 40) for _ in range(3000)] + [1, 2, (1 << 64) - 1], dtype=np.uint64)
    x = np.array([rng.randrange(-(1 << 40), 1 << 40) for _ in range(len(m))], dtype=np.int64)
    limbs = 40
    d = hipmem.DeviceArray((len(m), limbs), np.uint32).zero()
    engine.build_candidates_device(fam, m, x, limbs, d.ptr)
    got = d.to_host()
    for i in range(len(m)):
        assert pgap.from_limbs(got[i]) == int(m[i]) * K + int(x[i]), i


def test_build_candidates_rejects_overflow(engine):
    fam = engine.family(809, 1841070, 0)
    d = hipmem.DeviceArray((2, 35), np.uint32)
    with pytest.raises(pgap.PgapError) as ei:
        engine.build_candidates_device(fam, np.array([1 << 50, 1], np.uint64), np.array([0, 0], np.int64), 35,
                                       d.ptr)
    assert ei.value.status == pgap.PGAP_E_TOO_LARGE
    with pytest.raises(pgap.PgapError):
        engine.build_candidates_device(fam, np.array([0], np.uint64), np.array([-5], np.int64), 35, d.ptr)


@pytest.mark.parametrize("P,D,SL,ms", [(907, 2190, 11000, (1, 37, 199)), (503, 210, 4000, (1, 11, 1999))])
def test_filter_matches_oracle_sieve(engine, P, D, SL, ms):
    fam = engine.family(P, D, SL)
    sv = ol.Sieve(P, D, SL, 1_000_000)
    for mm in ms:
        xs = np.array([x for x in range(-SL, SL + 1) if x], dtype=np.int64)
        keep = engine.filter_candidates(fam, np.full(len(xs), mm, np.uint64), xs, 1_000_000)
        prev, nxt = sv(mm)
        want = set(-int(d) for d in prev) | set(int(d) for d in nxt)
        assert set(int(v) for v in xs[keep.astype(bool)]) == want


def test_filter_random_pairs_vs_trial_division(engine):
    fam = engine.family(13, 1, 0)
    K = 30030
    rng = random.Random(9)
    m = np.array([rng.randrange(1, 10 ** 9) for _ in range(5000)], dtype=np.uint64)
    x = np.array([rng.randrange(-10 ** 6, 10 ** 6) for _ in range(5000)], dtype=np.int64)
    keep = engine.filter_candidates(fam, m, x, 5000)
    primes = [p for p in range(2, 5001) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    for i in range(len(m)):
        n = int(m[i]) * K + int(x[i])
        want = all(n % p for p in primes)
        assert bool(keep[i]) == want, (int(m[i]), int(x[i]))
