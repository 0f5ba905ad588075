"""gap_test.py's probabilistic one-side skip: the oracle's restatement
(oracle.c or_*; gap_test_stats.py:241-284, 386-453, gap_test.py:184-202,
274-275) pinned by what the mathematics fixes -- an empty record set leaves
the product of (1 - p) over the counted candidates, a record set holding
every gap gives 1 (the terms telescope), hand-computed cases, closed forms of
the sieve and coprimality factors -- and then the product's side_prob.h
(through tests/shim/libside_prob.so) equal to the oracle, double for double.
The reference module imports gmpy2, which this image lacks, so it is not run."""
import ctypes
import math
import os
import random

import numpy as np
import pytest

import oracle_lib as ol

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(REPO, "tests", "shim", "libside_prob.so")

vp, u32, u64, i64, sz, dbl = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_size_t,
                              ctypes.c_double)


def oracle():
    lib = ol.load()
    lib.or_prob_prime_after_sieve.argtypes = [dbl, u64]
    lib.or_prob_prime_after_sieve.restype = dbl
    lib.or_setup_extended_gap.argtypes = [u32, u32, u32, dbl, vp, ctypes.c_int, vp]
    lib.or_setup_extended_gap.restype = dbl
    lib.or_prob_record_one_sided.argtypes = [vp, sz, i64, i64, vp, sz, dbl, u64, u64, u64, vp, sz, dbl]
    lib.or_prob_record_one_sided.restype = dbl
    lib.or_side_decisions.argtypes = [sz, vp, vp, dbl, vp]
    lib.or_side_decisions.restype = None
    return lib


def _ptr(a):
    return a.ctypes.data_as(vp)


def or_one_sided(records, megagap, prev_p, unknowns, pps, m, K_mod_d, d, ext, ppc):
    r = np.array(sorted(records), dtype=np.int64)
    u = np.array(unknowns, dtype=np.int32)
    x = np.array(ext, dtype=np.int32)
    return oracle().or_prob_record_one_sided(_ptr(r), len(r), megagap, prev_p, _ptr(u), len(u), pps, m, K_mod_d,
                                             d, _ptr(x), len(x), ppc)


def or_extended(SL, P, D, prob_prime):
    xs = np.zeros(2 * SL + 1, dtype=np.int32)
    n = ctypes.c_int()
    ppc = oracle().or_setup_extended_gap(SL, P, D, prob_prime, _ptr(xs), len(xs), ctypes.byref(n))
    return ppc, xs[:n.value].tolist()


def or_decisions(prob_record, new_prob, thr):
    a = np.array(prob_record, dtype=np.float64)
    b = np.array(new_prob, dtype=np.float64)
    out = np.zeros(len(a), dtype=np.uint8)
    oracle().or_side_decisions(len(a), _ptr(a), _ptr(b), thr, _ptr(out))
    return out.tolist()


# ---- the oracle, pinned ------------------------------------------------------

def test_hand_computed_three_unknowns():
    # prev 4, unknowns 2, 6, 10 -> gaps 6 (record), 10, 14 (record); p = 0.1:
    # 0.1 + 0.9 * 0.9 * 0.1 = 0.181, left 0.729; no extension -> 0.91
    assert or_one_sided({6, 14}, 0, 4, [2, 6, 10], 0.1, 1, 1, 2, [], 0.2) == pytest.approx(0.91, abs=1e-15)
    # extension x = 7, 11 with m * (K mod d) = 1, d = 2: 1 + x is even, so
    # gcd > 1 and both count at 0.2: gap 11 a record -> + 0.729 * 0.2; gap 15
    # not -> left 0.729 * 0.8 * 0.8 = 0.46656; total 0.181 + 0.1458 + 0.46656
    got = or_one_sided({6, 14, 11}, 0, 4, [2, 6, 10], 0.1, 1, 1, 2, [7, 11], 0.2)
    assert got == pytest.approx(0.79336, abs=1e-15)
    # d = 3: 1 + 7 = 8 and 1 + 11 = 12 -> only x = 11 (gcd 3) counts
    got = or_one_sided({6, 14, 11, 15}, 0, 4, [2, 6, 10], 0.1, 1, 1, 3, [7, 11], 0.2)
    assert got == pytest.approx(0.181 + 0.729 * 0.2 + 0.729 * 0.8, abs=1e-15)


def test_megagap_counts_gaps_above_it():
    # megagap 12: gaps 6 and 10 do not count, 14 and 15 (x = 11) do:
    # 0.81 * 0.1 + 0.729 * 0.2 + 0.729 * 0.8 = 0.81, the chance that neither
    # of the first two is prime
    got = or_one_sided(set(), 12, 4, [2, 6, 10], 0.1, 1, 1, 2, [11], 0.2)
    assert got == pytest.approx(0.81 * 0.1 + 0.729 * 0.2 + 0.729 * 0.8, abs=1e-15)
    assert got == pytest.approx(0.81, abs=1e-15)
    # megagap 14: only 15 counts
    got = or_one_sided(set(), 14, 4, [2, 6, 10], 0.1, 1, 1, 2, [11], 0.2)
    assert got == pytest.approx(0.729 * 0.2 + 0.729 * 0.8, abs=1e-15)


@pytest.mark.parametrize("seed", range(5))
def test_no_records_leaves_the_product_and_all_records_give_one(seed):
    rng = random.Random(seed)
    unknowns = sorted(rng.sample(range(1, 3000), 60))
    ext = sorted(rng.sample(range(3001, 6000), 200))
    pps, ppc, m, kd, d = rng.uniform(0.005, 0.05), rng.uniform(0.001, 0.02), rng.randrange(1, 10**6), 37, 2190
    counted = sum(1 for x in ext if math.gcd(m * kd + x, d) > 1)
    got = or_one_sided(set(), 0, 500, unknowns, pps, m, kd, d, ext, ppc)
    assert got == pytest.approx((1 - pps) ** len(unknowns) * (1 - ppc) ** counted, rel=1e-12)
    every = set(range(0, 10000))
    assert or_one_sided(every, 0, 500, unknowns, pps, m, kd, d, ext, ppc) == pytest.approx(1.0, abs=1e-12)


def test_extended_gap_closed_form():
    # P = 7, D = 6: K = 35, so the extension keeps x in (SL, 2SL] prime to 5
    # and 7; prob_prime_coprime = prob_prime / prod_{p <= 7} (1 - 1/p)
    # = prob_prime * 35 / 8
    ppc, xs = or_extended(50, 7, 6, 0.01)
    assert ppc == pytest.approx(0.01 * 35 / 8, rel=1e-15)
    assert xs == [x for x in range(51, 101) if math.gcd(x, 35) == 1]
    ppc, xs = or_extended(1000, 101, 2190, 0.001)
    K = ol.K(101, 2190)
    assert xs == [x for x in range(1001, 2001) if math.gcd(x, K) == 1]


def test_after_sieve_is_mertens():
    pp = 1 / 2000 - 1 / 2000 ** 2
    got = oracle().or_prob_prime_after_sieve(pp, 10**9)
    assert got == pytest.approx(pp * math.log(10**9) * math.exp(0.577215665), rel=1e-15)


def test_running_threshold_by_hand():
    # start 0 over 10 sides: threshold 0, 0.02 > 0 -> test (total 0.01, 12
    # sides); 0.01 / 12 = 8.3e-4 > 5e-4 -> skip (total 0.0195, 13 sides);
    # 0.0195 / 13 = 1.5e-3 < 1.6e-3 -> test (0.0295, 15); 0.0295 / 15 =
    # 1.97e-3 > 1e-3 -> skip
    got = or_decisions([0.01, 0.01, 0.01, 0.01], [0.02, 0.0005, 0.0016, 0.001], 0.0)
    assert got == [1, 0, 1, 0]
    # a prp-top-percent threshold of 1e-3 starts the average at 5e-4 per side
    assert or_decisions([0.01], [0.0004], 1e-3) == [0]
    assert or_decisions([0.01], [0.0006], 1e-3) == [1]


# ---- the product against the oracle -------------------------------------------

def product():
    if not os.path.exists(SHIM):
        pytest.skip("run __graft_entry__.build() first")
    lib = ctypes.CDLL(SHIM)
    lib.sp_new.argtypes = [vp, sz, i64, u32, u32, u32, u64, dbl, dbl, dbl]
    lib.sp_new.restype = vp
    lib.sp_free.argtypes = [vp]
    lib.sp_one_sided.argtypes = [vp, i64, vp, sz, u64]
    lib.sp_one_sided.restype = dbl
    lib.sp_decide.argtypes = [vp, dbl, dbl]
    lib.sp_prob_prime_coprime.argtypes = [vp]
    lib.sp_prob_prime_coprime.restype = dbl
    lib.sp_extended.argtypes = [vp, vp, sz]
    lib.sp_extended.restype = sz
    lib.sp_after_sieve.argtypes = [dbl, u64]
    lib.sp_after_sieve.restype = dbl
    lib.sp_prob_prime.argtypes = [dbl]
    lib.sp_prob_prime.restype = dbl
    return lib


@pytest.mark.parametrize("P,D,SL,megagap", [(907, 2190, 11000, 0), (1511, 2190, 15000, 0), (503, 210, 4000, 0),
                                            (907, 2190, 11000, 3000)])
def test_product_equals_oracle(P, D, SL, megagap):
    lib = product()
    rng = random.Random(P + megagap)
    log_n = ol.K_log(P, D) + math.log(1000)
    pp = lib.sp_prob_prime(log_n)
    assert pp == 1 / log_n - 1 / (log_n * log_n)
    pps = lib.sp_after_sieve(pp, 10**9)
    assert pps == oracle().or_prob_prime_after_sieve(pp, 10**9)
    records = sorted(rng.sample(range(2, int(35 * log_n), 2), int(10 * log_n)))
    K_mod_d = ol.K(P, D) % D
    thr = rng.choice([0.0, 2e-4])
    r = np.array(records, dtype=np.int64)
    h = lib.sp_new(_ptr(r), len(r), megagap, P, D, SL, K_mod_d, pp, pps, thr)
    try:
        ppc, ext = or_extended(SL, P, D, pp)
        assert lib.sp_prob_prime_coprime(h) == ppc
        out = np.zeros(len(ext) + 1, dtype=np.int32)
        assert lib.sp_extended(h, _ptr(out), len(out)) == len(ext)
        assert out[:len(ext)].tolist() == ext
        probs, news = [], []
        for m in rng.sample(range(1, 10**6), 40):
            if math.gcd(m, D) != 1:
                continue
            prev_p = rng.randrange(1, 3 * SL)
            nxt = np.array(sorted(rng.sample(range(1, SL + 1), rng.randrange(0, 200))), dtype=np.int32)
            got = lib.sp_one_sided(h, prev_p, _ptr(nxt), len(nxt), m)
            want = or_one_sided(records, megagap, prev_p, nxt.tolist(), pps, m, K_mod_d, D, ext, ppc)
            assert got == want, (m, prev_p)
            probs.append(float(np.float32(rng.uniform(0, 4e-4))))
            news.append(got)
        want = or_decisions(probs, news, thr)
        assert [lib.sp_decide(h, a, b) for a, b in zip(probs, news)] == want
    finally:
        lib.sp_free(h)


def expected_decisions(P, D, SL, max_prime, mstart, rows, records, megagap, prob_record, prob_threshold=0.0):
    """The oracle's one-side decisions for m's in order: rows = [(m, prev_p,
    next unknowns)], prob_record[i] = m_stats' value for rows[i]; ln N taken
    at mstart (gap_test.py:515)."""
    M_log = ol.K_log(P, D) + math.log(mstart)
    pp = 1 / M_log - 1 / (M_log * M_log)
    pps = oracle().or_prob_prime_after_sieve(pp, max_prime)
    ppc, ext = or_extended(SL, P, D, pp)
    kd = ol.K(P, D) % D
    news = [or_one_sided(records, megagap, prev_p, list(nxt), pps, m, kd, D, ext, ppc) for m, prev_p, nxt in rows]
    return or_decisions(prob_record, news, prob_threshold)
