"""Mixed widths in one persistent launch: verdicts against the oracle and
against the per-width launches of the same numbers."""
import math
import random

import numpy as np
import pytest

import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu


def mix(count, seed):
    """The uneven-batching mix: 40% 503#/210, 40% 1511#/2190, 20%
    809#/1841070, m ~ U[1, 1e6) coprime to d, x ~ U[-20000, 20000]."""
    rng = random.Random(seed)
    fams = [(503, 210, 0.4), (1511, 2190, 0.4), (809, 1841070, 0.2)]
    Ks = {(P, D): ol.K(P, D) for P, D, _ in fams}
    vals = []
    while len(vals) < count:
        r = rng.random()
        P, D, _ = fams[0] if r < 0.4 else fams[1] if r < 0.8 else fams[2]
        m = rng.randrange(1, 1_000_000)
        if math.gcd(m, D) != 1:
            continue
        vals.append(m * Ks[(P, D)] + rng.randrange(-20000, 20001))
    return vals


def limb_rows(vals):
    return [pgap.to_limbs([v], pgap.limbs_needed(v))[0] for v in vals]


def test_mixed_matches_oracle_and_per_width(engine):
    vals = mix(3000, 21) + [341, 561, 2047, (1 << 521) - 1, (1 << 2127) - 1, (1 << 1279) - 1, 0, 1, 2, 3]
    rows = limb_rows(vals)
    got = engine.prp_mixed(rows)
    for width in sorted(set(len(r) for r in rows)):
        idx = [i for i, r in enumerate(rows) if len(r) == width]
        arr = np.stack([rows[i] for i in idx])
        want = ol.verdicts(arr, 0)
        assert np.array_equal(got[idx], want), width
        assert np.array_equal(engine.prp(arr), want), width
    assert got.sum() > 0


def test_mixed_sprp(engine):
    vals = mix(600, 22) + [2047, 3277, 25326001, (1 << 607) - 1]
    rows = limb_rows(vals)
    got = engine.prp_mixed(rows, pgap.SPRP, (2, 3))
    for width in sorted(set(len(r) for r in rows)):
        idx = [i for i, r in enumerate(rows) if len(r) == width]
        want = ol.verdicts(np.stack([rows[i] for i in idx]), 1, (2, 3))
        assert np.array_equal(got[idx], want)


def test_mixed_rejects_too_wide(engine):
    # the widest bucket: 8 lanes x 20 digits of 28 bits, less 4 bits of
    # headroom for Fermat, 8 for the strong test
    engine.prp_mixed(limb_rows([3, (1 << 4475) + 1]))
    with pytest.raises(pgap.PgapError) as ei:
        engine.prp_mixed(limb_rows([3, (1 << 4476) + 1]))
    assert ei.value.status == pgap.PGAP_E_TOO_LARGE
    engine.prp_mixed(limb_rows([3, (1 << 4471) + 1]), pgap.SPRP, (2,))
    with pytest.raises(pgap.PgapError) as ei:
        engine.prp_mixed(limb_rows([3, (1 << 4472) + 1]), pgap.SPRP, (2,))
    assert ei.value.status == pgap.PGAP_E_TOO_LARGE


# Mersenne primes across the lane-group buckets (2^p - 1 for p = 2203, 2281,
# 3217, 4253, 4423), composites next to them, and candidates of the large
# families (1511#/2190: 2128 bits; 3001#/2190: 4231-4247 bits)
MERSENNE = [2203, 2281, 3217, 4253, 4423]


def wide_mix(count, seed):
    rng = random.Random(seed)
    vals = []
    K15, K30 = ol.K(1511, 2190), ol.K(3001, 2190)
    for bits in (2182, 2500, 2684, 2700, 3000, 3356, 3400, 3580, 3600, 4000, 4252, 4300, 4476):
        for _ in range(8):
            vals.append(rng.getrandbits(bits) | (1 << (bits - 1)) | 1)
    while len(vals) < count:
        m = rng.randrange(1, 100_000)
        if math.gcd(m, 2190) != 1:
            continue
        K = K30 if rng.random() < 0.6 else K15
        vals.append(m * K + rng.randrange(-40000, 40001))
    for p in MERSENNE:
        vals += [(1 << p) - 1, (1 << p) + 1, ((1 << p) - 1) * 3]
    return vals


@pytest.mark.parametrize("kind,bases", [(pgap.FERMAT2, ()), (pgap.SPRP, (2, 3))])
def test_mixed_lane_group_buckets(engine, kind, bases):
    vals = wide_mix(700, 23) + mix(300, 24)
    if kind == pgap.SPRP:
        vals = [v for v in vals if v.bit_length() <= 4472]
    rows = limb_rows(vals)
    got = engine.prp_mixed(rows, kind, bases) if bases else engine.prp_mixed(rows)
    # the waves of one workgroup run buckets of different widths at once:
    # repeated launches place the tiles differently
    for _ in range(3):
        again = engine.prp_mixed(rows, kind, bases) if bases else engine.prp_mixed(rows)
        assert np.array_equal(again, got)
    okind = 0 if kind == pgap.FERMAT2 else 1
    for width in sorted(set(len(r) for r in rows)):
        idx = [i for i, r in enumerate(rows) if len(r) == width]
        arr = np.stack([rows[i] for i in idx])
        want = ol.verdicts(arr, okind, bases) if bases else ol.verdicts(arr, okind)
        assert np.array_equal(got[idx], want), width
    for p in MERSENNE:
        i = vals.index((1 << p) - 1)
        assert got[i] == 1 and got[i + 2] == 0, p
    # the same numbers in the per-width batch path
    wide = [i for i, v in enumerate(vals) if v.bit_length() > 4000]
    arr = np.stack([pgap.to_limbs([vals[i]], 140)[0] for i in wide])
    per = engine.prp(arr, kind, bases) if bases else engine.prp(arr)
    assert np.array_equal(got[wide], per)
