"""The reference's CPU-runnable case end to end on the GPU: P = 503, d = 210,
sieve length 4000, every m in [1, 2000) coprime to 210 (458 m's), unknowns
sieved to 1e8 on the device, scanned outward with Baillie-PSW confirmation.
Every row equals the oracle's test_interval_cpu row (gap_test_common.cpp:
128-181); the CLI pair gap_sieve_hip -> gap_test_hip prints what
gap_test_simple would (gap_test_simple.cpp:101-200) on the written file."""
import math
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import pgap
from test_gpu_cli import CLI, expected_stdout, normalise
from test_gpu_sieve import SIEVE

pytestmark = pytest.mark.gpu

P, D, SL, M_HI, LIMIT = 503, 210, 4000, 2000, 100_000_000
MS = [m for m in range(1, M_HI) if math.gcd(m, D) == 1]


def test_every_row_matches_the_oracle():
    assert len(MS) == 458
    with pgap.Engine(0) as e:
        fam = e.family(P, D, SL)
        lists = e.sieve(fam, np.array(MS, dtype=np.uint64), LIMIT, SL)
        for m, (p, q) in zip(MS, lists):
            assert e.scan_submit(fam, m, p, q, tag=m) == pgap.PGAP_OK
        rows = {}
        while e.scan_in_flight():
            for r in e.scan_poll(wait=True):
                rows[r.m] = r
    want = ol.scan_many(P, D, SL, MS, lists, kind=0)
    tests = 0
    for i, m in enumerate(MS):
        r = rows[m]
        w = ol.test_interval(P, D, m, SL, *lists[i], kind=0) if i % 50 == 0 else None
        assert (r.prev_p, r.next_p) == (want[i][0], want[i][1]), m
        assert r.p_tests + r.n_tests == want[i][2], m
        if w:
            assert (r.p_tests, r.n_tests) == (w["p_tests"], w["n_tests"]), m
        tests += r.p_tests + r.n_tests
    # about 2 / prob_prime_after_sieve tests per m (gap_common.cpp:581-601): ~29
    assert 20 < tests / len(MS) < 40
    # a few sieve lists against the oracle sieve
    sv = ol.Sieve(P, D, SL, LIMIT)
    for i in (0, 200, 457):
        p, q = sv(MS[i])
        assert np.array_equal(p, lists[i][0]) and np.array_equal(q, lists[i][1]), MS[i]


def test_cli_on_the_device_sieved_file(tmp_path):
    r = subprocess.run([SIEVE, "-p", str(P), "-d", str(D), "--mstart", "1", "--minc", str(M_HI), "--max-prime",
                        "100", "--sieve-length", str(SL), "--save", "-qq"], cwd=tmp_path, capture_output=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    path = tmp_path / "unknowns" / f"{P}_{D}_1_{M_HI}_s{SL}_l100M.txt"
    lines = path.read_bytes().splitlines(keepends=True)
    assert len(lines) == len(MS)
    r = subprocess.run([CLI, "-u", str(path), "--min-merit", "5", "-q"], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    assert normalise(r.stdout.decode()) == expected_stdout(P, D, 1, M_HI, SL, lines, 5.0, 1)
