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
    with pytest.raises(pgap.PgapError) as ei:
        engine.prp_mixed(limb_rows([3, (1 << 3000) + 1]))
    assert ei.value.status == pgap.PGAP_E_TOO_LARGE
