"""Rows at the configured sieve depths of the two large families:
1511#/2190 (SL 40000, unknowns sieved to 1e9: 2128-bit numbers) and
3001#/2190 (SL 40000, sieved to 1e10: 4231-4248 bits), m's drawn as the
serving bench draws them.  Every scan row equals the oracle's
test_interval_cpu row (gap_test_common.cpp:128-181) on the same lists,
including the search beyond the sieve length where a gap needs it; at 1e9
the device sieve's lists also equal the oracle sieve's (at 1e10 the CPU
sieve would take minutes, so the lists there are checked by the sieve's
own md5 tests at smaller depths and by the rows)."""
import math
import random

import numpy as np
import pytest

import oracle_lib as ol
import pgap
from test_gpu_scan import run_scan

pytestmark = pytest.mark.gpu

SL = 40000


def draw_ms(D, m_hi, count, seed):
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        m = rng.randrange(1, m_hi)
        if math.gcd(m, D) == 1 and m not in out:
            out.append(m)
    return out


def scan_and_compare(P, D, limit, ms):
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        lists = e.sieve(fam, np.array(ms, dtype=np.uint64), limit, SL)
        rows = run_scan(e, fam, [(m, p, q) for m, (p, q) in zip(ms, lists)])
    want = ol.scan_many(P, D, SL, ms, lists, kind=0)
    for i, m in enumerate(ms):
        r = rows[m]
        assert (r.prev_p, r.next_p) == (want[i][0], want[i][1]), m
        assert r.p_tests + r.n_tests == want[i][2], m
    w = ol.test_interval(P, D, ms[0], SL, *lists[0], kind=0)
    r = rows[ms[0]]
    assert (r.p_tests, r.n_tests) == (w["p_tests"], w["n_tests"])
    return lists


def test_1511_family_sieved_to_1e9():
    P, D, LIMIT = 1511, 2190, 10 ** 9
    ms = draw_ms(D, 200_000, 24, 1511)
    lists = scan_and_compare(P, D, LIMIT, ms)
    # ~1715 unknowns per m at this depth (the serving bench's figure)
    assert 1500 < np.mean([len(p) + len(q) for p, q in lists]) < 1950
    sv = ol.Sieve(P, D, SL, LIMIT)
    try:
        for i in (0, 11, 23):
            p, q = sv(ms[i])
            assert np.array_equal(p, lists[i][0]) and np.array_equal(q, lists[i][1]), ms[i]
    finally:
        sv.close()


def test_3001_family_sieved_to_1e10():
    P, D, LIMIT = 3001, 2190, 10 ** 10
    ms = draw_ms(D, 100_000, 6, 3001)
    lists = scan_and_compare(P, D, LIMIT, ms)
    # ~1162 unknowns per m at this depth
    assert 1000 < np.mean([len(p) + len(q) for p, q in lists]) < 1350
