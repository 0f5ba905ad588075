"""The test_runner_t drop-in on the GPU, driven on mpz_t inputs as
run_gpu_thread drives the reference's (gap_test_gpu.cu:236-248): rounds = 1
is a strong test to base 2 (generate_primes(1) = {2}, miller_rabin.h:269-286,
317-319), so the base-2 Fermat pseudoprimes 341 and 561 fail it while the
strong base-2 pseudoprimes 2047 and 3277 pass; rounds = 2 tests bases 2 and
3.  results[i] against the oracle's strong tests."""
import os
import random
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "tests", "shim", "run_test_main")


def run(rounds, vals):
    r = subprocess.run([DRIVER, str(rounds)], input="".join(f"{v:x}\n" for v in vals).encode(),
                       capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    return [int(x) for x in r.stdout.split()]


def test_rounds_follow_generate_primes():
    rng = random.Random(1108)
    K = ol.K(809, 1841070)
    vals = [341, 561, 645, 1105, 2047, 3277, 4033, 4681, 8321, 1373653, 25326001, 3215031751]
    vals += [(1 << 1109) - 1, (1 << 1279) - 1]
    while len(vals) < 400:
        v = rng.randrange(10_000_000, 12_000_000) * K + rng.randrange(-15000, 15001)
        vals.append(v | 1)
    x = vals[-1]
    while not ol.probab_prime(x):
        x += 2
    vals.append(x)
    got1 = run(1, vals)
    assert got1 == [ol.sprp(v, 2) for v in vals]
    assert got1[:6] == [0, 0, 0, 0, 1, 1]  # 341, 561, 645, 1105 fail; 2047, 3277 pass
    got2 = run(2, vals)
    assert got2 == [int(ol.sprp(v, 2) and ol.sprp(v, 3)) for v in vals]
    assert got2[vals.index(1373653)] == 1 and got2[vals.index(2047)] == 0  # 1373653: strong to bases 2 and 3
    got3 = run(3, vals)
    assert got3 == [int(all(ol.sprp(v, b) for b in (2, 3, 5))) for v in vals]
    assert got3[vals.index(25326001)] == 1 and got3[-1] == 1
