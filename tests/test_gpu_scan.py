"""Streaming scan parity: per-m rows from the device-resident scan against
the oracle's restatement of test_interval_cpu (gap_test_common.cpp:128-181),
which is itself pinned to the reference README rows."""
import math

import numpy as np
import pytest

import oracle_lib as ol
import pgap
import unknown_format as uf

pytestmark = pytest.mark.gpu


def lines_907(unknowns_907):
    return [uf.parse_line(ln) for ln in unknowns_907.splitlines(keepends=True)]


def run_scan(engine, fam, reqs, depth=None, confirm=None):
    """Submit everything (retrying on admission back-pressure), collect rows."""
    rows = {}
    pending = list(reqs)
    while pending or engine.scan_in_flight():
        while pending:
            m, p, n = pending[0]
            st = engine.scan_submit(fam, m, np.array(p, np.int32), np.array(n, np.int32), tag=m)
            if st == pgap.PGAP_E_OOM:
                break
            pending.pop(0)
        for r in engine.scan_poll(wait=True):
            assert r.m not in rows
            rows[r.m] = r
    return rows


def test_scan_rows_907_match_oracle(unknowns_907):
    lines = lines_907(unknowns_907)
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 11000)
        rows = run_scan(e, fam, lines)
    assert len(rows) == len(lines)
    total = 0
    for m, p, n in lines:
        want = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        r = rows[m]
        assert (r.prev_p, r.next_p) == (want["prev_p"], want["next_p"]), m
        assert (r.p_tests, r.n_tests) == (want["p_tests"], want["n_tests"]), m
        assert r.flags == 0 and r.tag == m
        total += r.p_tests + r.n_tests
    assert total == 2590  # README.md:241-262 prp-test total over the 53 m's


# oracle in-sieve test per confirmation: Fermat base 2, strong base 2
# (the reference GPU tester), mpz_probab_prime_p (gap_test_simple)
ORACLE_KIND = {pgap.CONFIRM_NONE: 1, pgap.CONFIRM_SPRP2: 2, pgap.CONFIRM_BPSW: 0}


@pytest.mark.parametrize("depth,confirm", [(1, pgap.CONFIRM_NONE), (3, pgap.CONFIRM_SPRP2),
                                           (16, pgap.CONFIRM_BPSW)])
def test_scan_depth_and_confirmation_do_not_change_rows(unknowns_907, depth, confirm):
    lines = lines_907(unknowns_907)[:20]
    with pgap.Engine(0) as e:
        e.scan_config(depth, confirm)
        fam = e.family(907, 2190, 11000)
        rows = run_scan(e, fam, lines)
    for m, p, n in lines:
        want = ol.test_interval(907, 2190, m, 11000, p, n, kind=ORACLE_KIND[confirm])
        assert (rows[m].prev_p, rows[m].next_p, rows[m].p_tests, rows[m].n_tests) == \
            (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"])


def test_alone_equals_mixed(unknowns_907):
    lines = lines_907(unknowns_907)
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 11000)
        mixed = run_scan(e, fam, lines)
        alone = {}
        for m, p, n in lines[::7]:
            alone.update(run_scan(e, fam, [(m, p, n)]))
    for m, r in alone.items():
        a = (r.prev_p, r.next_p, r.p_tests, r.n_tests, r.flags)
        b = (mixed[m].prev_p, mixed[m].next_p, mixed[m].p_tests, mixed[m].n_tests, mixed[m].flags)
        assert a == b, m


def test_alone_equals_mixed_across_round_variants():
    # 1511#/2190 (2128 bits): 640 m's in flight are 1280 scanning sides, above
    # the latency-first threshold, so the mix starts on one-lane rounds and
    # ends on 8-lane ones; an m alone runs on 8-lane rounds throughout.  Rows
    # must not depend on the variant a round used.
    P, D, SL = 1511, 2190, 40000
    ms = [m for m in range(1, 3000) if math.gcd(m, D) == 1][:640]
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        lists = e.sieve(fam, np.array(ms, dtype=np.uint64), 100_000_000, SL)
        reqs = [(m, p, q) for m, (p, q) in zip(ms, lists)]
        mixed = run_scan(e, fam, reqs)
        alone = {}
        for req in reqs[::97]:
            alone.update(run_scan(e, fam, [req]))
    assert len(mixed) == len(ms) and len(alone) == 7
    for m, r in alone.items():
        a = (r.prev_p, r.next_p, r.p_tests, r.n_tests, r.flags)
        b = (mixed[m].prev_p, mixed[m].next_p, mixed[m].p_tests, mixed[m].n_tests, mixed[m].flags)
        assert a == b, m
    # spot-check the oracle on two of them
    for m, p, q in reqs[:2]:
        w = ol.test_interval(P, D, m, SL, p, q, kind=0)
        assert (mixed[m].prev_p, mixed[m].next_p) == (w["prev_p"], w["next_p"]), m


@pytest.mark.parametrize("confirm", [pgap.CONFIRM_NONE, pgap.CONFIRM_SPRP2, pgap.CONFIRM_BPSW])
def test_confirmation_rejects_fermat_pseudoprimes(confirm):
    # K = 11#/1155 = 2 (init_K asserts K > 1), so N = 2m -+ x is small enough
    # to place base-2 Fermat pseudoprimes in the lists: 404 - 63 = 341 =
    # 11*31 and 404 + 157 = 561 = 3*11*17 (a Carmichael number).  The
    # Fermat-only scan stops on them; neither is a strong base-2 PRP, so the
    # reference GPU tester's test (strong base 2, ROUNDS=1) and Baillie-PSW
    # (mpz_probab_prime_p) both resume past them.
    P, D, SL, m = 11, 1155, 200, 202
    prev = [63, 65, 67, 71]    # 341 (psp), 339 = 3*113, 337 prime, 333
    nxt = [157, 159, 165]      # 561 (Carmichael), 563 prime, 569 prime
    want = ol.test_interval(P, D, m, SL, prev, nxt, kind=ORACLE_KIND[confirm])
    assert (want["prev_p"], want["next_p"]) == ((63, 157) if confirm == pgap.CONFIRM_NONE else (67, 159))
    with pgap.Engine(0) as e:
        e.scan_config(4, confirm)
        fam = e.family(P, D, SL)
        rows = run_scan(e, fam, [(m, prev, nxt)])
    r = rows[m]
    assert (r.prev_p, r.next_p, r.p_tests, r.n_tests, r.flags) == \
        (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"], 0)


SPSP_235 = [25326001, 3215031751, 2152302898747, 3474749660383, 341550071728321]


def test_bpsw_confirmation_rejects_strong_pseudoprimes():
    # Strong pseudoprimes to bases 2, 3, 5 (OEIS A014233 terms) placed first
    # on a side: K = 2, centre N - 1 (next side, x = 1) or N + 1 (prev side).
    # The strong base-2 confirmation (the reference GPU tester's verdict)
    # stops on them; Baillie-PSW rejects them as mpz_probab_prime_p does, and
    # the side resumes to the next prime.
    P, D, SL = 11, 1155, 200
    odd = list(range(1, SL, 2))
    reqs, psp_side = [], {}
    for n in SPSP_235:
        reqs.append(((n - 1) // 2, odd, odd))
        psp_side[(n - 1) // 2] = "next_p"
        reqs.append(((n + 1) // 2, odd, odd))
        psp_side[(n + 1) // 2] = "prev_p"
    rows = {}
    for confirm in (pgap.CONFIRM_SPRP2, pgap.CONFIRM_BPSW):
        with pgap.Engine(0) as e:
            e.scan_config(64, confirm)
            fam = e.family(P, D, SL)
            rows[confirm] = run_scan(e, fam, reqs)
    for m, p, q in reqs:
        for confirm, kind in ((pgap.CONFIRM_SPRP2, 2), (pgap.CONFIRM_BPSW, 0)):
            w = ol.test_interval(P, D, m, SL, p, q, kind=kind)
            r = rows[confirm][m]
            assert (r.prev_p, r.next_p, r.p_tests, r.n_tests) == \
                (w["prev_p"], w["next_p"], w["p_tests"], w["n_tests"]), (m, confirm)
        # the pseudoprime's side: distance 1 under the strong base-2 test only
        side = psp_side[m]
        assert getattr(rows[pgap.CONFIRM_SPRP2][m], side) == 1, m
        assert getattr(rows[pgap.CONFIRM_BPSW][m], side) > 1, m


def test_small_centres_fall_back_to_small_primes():
    # K = 2, empty unknown lists: both sides fall back beyond SL.  Below
    # 2^14 the fallback screen keeps N that are themselves small primes:
    # m = 4 (centre 8, SL 4): prevprime(4) = 3, nextprime(12) = 13;
    # m = 150 (centre 300, SL 100): 199 and 401.
    P, D = 11, 1155
    for m, SL, want in ((4, 4, (5, 5)), (150, 100, (101, 101))):
        w = ol.test_interval(P, D, m, SL, [], [], kind=0)
        assert (w["prev_p"], w["next_p"]) == want
        with pgap.Engine(0) as e:
            fam = e.family(P, D, SL)
            r = run_scan(e, fam, [(m, [], [])])[m]
        assert (r.prev_p, r.next_p) == want, m
        assert r.flags == pgap.ROW_PREV_FALLBACK | pgap.ROW_NEXT_FALLBACK


@pytest.mark.parametrize("confirm", [pgap.CONFIRM_NONE, pgap.CONFIRM_BPSW])
def test_width_grows_past_a_power_of_two(confirm):
    # K = 2 and centres 2m just below 2^48: the next-side candidates 2m + x
    # reach 49 bits (the next primes of the first two m's are 2^48 + 21), one
    # bit above the widest N the 2-digit strong-PRP variant registers
    # (28*2 - 8).  A round's width counts the room above each centre
    # (20, 6 and 1400 here), so these rounds run on a wider variant; rows
    # match the oracle.
    P, D, SL = 11, 1155, 2000
    ms = [2**47 - 10, 2**47 - 3, 2**47 - 700]
    sv = ol.Sieve(P, D, SL, 10000)
    lists = {m: sv(m) for m in ms}
    kind = 0 if confirm else 1
    want = {m: ol.test_interval(P, D, m, SL, *lists[m], kind=kind) for m in ms}
    assert (2 * ms[0] + want[ms[0]]["next_p"]).bit_length() == 49
    with pgap.Engine(0) as e:
        e.scan_config(64, confirm)
        fam = e.family(P, D, SL)
        rows = run_scan(e, fam, [(m, *lists[m]) for m in ms])
    for m in ms:
        r, w = rows[m], want[m]
        assert (r.prev_p, r.next_p, r.p_tests, r.n_tests) == (w["prev_p"], w["next_p"], w["p_tests"], w["n_tests"]), m


@pytest.mark.parametrize("min_merit", [6.0, 7.5])
def test_side_skip_follows_gap_test_gpu_rule(unknowns_907, min_merit):
    # gap_test_gpu.cu:591-617: with the prev PRP found and the next side not
    # started, (float)(prev_p / (K_log + log m)) < min_merit/2 - 2 ends the m.
    lines = lines_907(unknowns_907)
    K_log = ol.K_log(907, 2190)
    thr = np.float32(np.float32(min_merit) / np.float32(2) - np.float32(2))
    with pgap.Engine(0) as e:
        e.scan_side_skip(True, min_merit)
        fam = e.family(907, 2190, 11000)
        rows = run_scan(e, fam, lines)
    skipped = 0
    for m, p, n in lines:
        want = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        r = rows[m]
        skip = np.float32(want["prev_p"] / (K_log + math.log(m))) < thr
        skipped += bool(skip)
        if skip:
            expect = (want["prev_p"], 0, want["p_tests"], 0, pgap.ROW_NEXT_SKIPPED)
        else:
            expect = (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"], 0)
        assert (r.prev_p, r.next_p, r.p_tests, r.n_tests, r.flags) == expect, m
    assert 0 < skipped < len(lines)


def test_fallback_beyond_sieve_length():
    # README.md:40-46 record: 11244911 * 809#/1841070, next prime +17354 is
    # beyond SL = 15000, so the next side runs out of unknowns and continues
    # past SL (gap_test_common.cpp:170-178).
    m, P, D, SL = 11244911, 809, 1841070, 15000
    sv = ol.Sieve(P, D, SL, 1_000_000)
    prev, nxt = sv(m)
    want = ol.test_interval(P, D, m, SL, prev, nxt, kind=0)
    assert (want["prev_p"], want["next_p"], want["next_fallback"]) == (5788, 17354, 1)
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        rows = run_scan(e, fam, [(m, prev, nxt)])
        r = rows[m]
        assert (r.prev_p, r.next_p) == (5788, 17354)
        assert r.flags == pgap.ROW_NEXT_FALLBACK
        assert (r.p_tests, r.n_tests) == (want["p_tests"], want["n_tests"])
        # empty lists on both sides: both fall back
        fam2 = e.family(P, D, 5000)
        rows = run_scan(e, fam2, [(m, [], [])])
        r = rows[m]
        assert (r.prev_p, r.next_p, r.p_tests, r.n_tests) == (5788, 17354, 0, 0)
        assert r.flags == pgap.ROW_PREV_FALLBACK | pgap.ROW_NEXT_FALLBACK


def test_scan_2100_bit_family_small():
    # 1511#/2190 (the 2128-bit serving family) on a few m's sieved to 1e6
    P, D, SL = 1511, 2190, 40000
    sv = ol.Sieve(P, D, SL, 1_000_000)
    ms = [m for m in range(1, 60) if math.gcd(m, D) == 1][:6]
    reqs = [(m,) + sv(m) for m in ms]
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        rows = run_scan(e, fam, reqs)
    for m, p, n in reqs:
        want = ol.test_interval(P, D, m, SL, p, n, kind=0)
        r = rows[m]
        assert (r.prev_p, r.next_p, r.p_tests, r.n_tests) == \
            (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"]), m


@pytest.mark.parametrize("lanes", [0, 4, 8, 16])
def test_scan_4300_bit_family_small(lanes):
    # 3001#/2190 (N ~ 4231-4248 bits): lane-group variants in the scan (with
    # lanes=0 two m's take latency-first rounds on the 16-lane variants;
    # lanes=4 runs them on the 4 x 38 quads that load rounds use)
    P, D, SL = 3001, 2190, 40000
    sv = ol.Sieve(P, D, SL, 1_000_000)
    reqs = [(m,) + sv(m) for m in (1, 7)]
    with pgap.Engine(0, lanes=lanes) as e:
        fam = e.family(P, D, SL)
        rows = run_scan(e, fam, reqs)
    for m, p, n in reqs:
        want = ol.test_interval(P, D, m, SL, p, n, kind=0)
        r = rows[m]
        assert (r.prev_p, r.next_p, r.p_tests, r.n_tests) == \
            (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"]), m


def test_budget_and_admission(unknowns_907):
    lines = lines_907(unknowns_907)
    with pgap.Engine(0, hbm_budget=1000) as e:
        fam = e.family(907, 2190, 11000)
        with pytest.raises(pgap.PgapError) as ei:
            e.scan_submit(fam, 1, np.array([2], np.int32), np.array([4], np.int32))
        assert ei.value.status == pgap.PGAP_E_TOO_LARGE
    # room for a handful of m's: submissions beyond the backlog are refused
    # with PGAP_E_OOM and succeed after polling
    with pgap.Engine(0, hbm_budget=3 << 20) as e:
        fam = e.family(907, 2190, 11000)
        refused = 0
        rows = {}
        pending = list(lines)
        while pending or e.scan_in_flight():
            while pending:
                m, p, n = pending[0]
                if e.scan_submit(fam, m, np.array(p, np.int32), np.array(n, np.int32)) == pgap.PGAP_E_OOM:
                    refused += 1
                    break
                pending.pop(0)
            for r in e.scan_poll(wait=True):
                rows[r.m] = r
        assert refused > 0 and len(rows) == len(lines)
        for m, p, n in lines[:10]:
            want = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
            assert (rows[m].prev_p, rows[m].next_p) == (want["prev_p"], want["next_p"])


def test_submit_validates_lists():
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 11000)
        tiny = e.family(11, 1155, 200)
        # centre m*K = 200 below SL + 4: the prev side would reach N < 4
        with pytest.raises(pgap.PgapError) as ei:
            e.scan_submit(tiny, 100, np.array([1], np.int32), np.array([1], np.int32))
        assert ei.value.status == pgap.PGAP_E_ARG
        with pytest.raises(pgap.PgapError):
            e.scan_submit(fam, 1, np.array([0], np.int32), np.array([], np.int32))
        with pytest.raises(pgap.PgapError):
            e.scan_submit(fam, 1, np.array([11001], np.int32), np.array([], np.int32))
        with pytest.raises(pgap.PgapError):
            e.scan_submit(fam + 5, 1, np.array([2], np.int32), np.array([], np.int32))


def test_late_family_wider_than_the_slots_is_refused():
    # The slots' unknown lists are sized at the first submit from the
    # families registered by then; a later family with a longer sieve
    # length is refused before anything is launched for it.
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 5000)
        assert e.scan_submit(fam, 1, np.array([4], np.int32), np.array([2], np.int32)) == pgap.PGAP_OK
        big = e.family(907, 2190, 30000)
        with pytest.raises(pgap.PgapError) as ei:
            e.scan_submit(big, 1, np.array(list(range(2, 25000, 2)), np.int32), np.array([2], np.int32))
        assert ei.value.status == pgap.PGAP_E_STATE
        while e.scan_in_flight():
            e.scan_poll(wait=True)


def test_poll_timeout_then_row():
    # a 3001#/2190 m takes tens of milliseconds: a 1 ms poll right after the
    # submit gives PGAP_E_TIMEOUT (an empty list), the row comes later
    P, D, SL = 3001, 2190, 40000
    sv = ol.Sieve(P, D, SL, 1_000_000)
    p, n = sv(1)
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        assert e.scan_submit(fam, 1, p, n, tag=1) == pgap.PGAP_OK
        assert e.scan_poll(wait=True, timeout_ms=1) == []
        assert "no row within 1 ms" in e.last_error()
        rows = []
        while e.scan_in_flight():
            rows += e.scan_poll(wait=True, timeout_ms=60_000)
    want = ol.test_interval(P, D, 1, SL, p, n, kind=0)
    assert [(r.m, r.prev_p, r.next_p) for r in rows] == [(1, want["prev_p"], want["next_p"])]


def test_one_sided_requests_return_their_half(unknowns_907):
    # PGAP_SIDES_PREV / PGAP_SIDES_NEXT (the two halves gap_test.py's
    # one-side skip scans, gap_test.py:169-202): each m submitted three times
    # (both, prev only, next only) under distinct tags, all in flight together
    lines = lines_907(unknowns_907)
    rows = {}
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 11000)
        todo = [(m, p, n, s) for m, p, n in lines for s in (pgap.SIDES_BOTH, pgap.SIDES_PREV, pgap.SIDES_NEXT)]
        while todo or e.scan_in_flight():
            while todo:
                m, p, n, s = todo[0]
                if e.scan_submit(fam, m, np.array(p, np.int32), np.array(n, np.int32), tag=4 * m + s,
                                 sides=s) == pgap.PGAP_E_OOM:
                    break
                todo.pop(0)
            for r in e.scan_poll(wait=True):
                rows[r.tag] = r
    assert len(rows) == 3 * len(lines)
    for m, p, n in lines:
        want = ol.test_interval(907, 2190, m, 11000, p, n, kind=0)
        both, prev, nxt = (rows[4 * m + s] for s in (pgap.SIDES_BOTH, pgap.SIDES_PREV, pgap.SIDES_NEXT))
        assert (both.prev_p, both.next_p, both.p_tests, both.n_tests) == \
            (want["prev_p"], want["next_p"], want["p_tests"], want["n_tests"]), m
        assert (prev.prev_p, prev.next_p, prev.p_tests, prev.n_tests) == (want["prev_p"], 0, want["p_tests"], 0), m
        assert (nxt.prev_p, nxt.next_p, nxt.p_tests, nxt.n_tests) == (0, want["next_p"], 0, want["n_tests"]), m
        assert prev.flags == pgap.ROW_NEXT_SKIPPED and nxt.flags == pgap.ROW_PREV_SKIPPED and both.flags == 0
    with pgap.Engine(0) as e:
        fam = e.family(907, 2190, 11000)
        m, p, n = lines[0]
        with pytest.raises(pgap.PgapError):
            e.scan_submit(fam, m, np.array(p, np.int32), np.array(n, np.int32), sides=3)
