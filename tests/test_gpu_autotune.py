"""Autotuner: the table it writes covers every requested width and every
test (Fermat, strong PRP, strong Lucas), and an engine created from that
table gives the oracle's verdicts (the tuning changes speed only, never
results)."""
import random

import numpy as np
import pytest

import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu

WIDTHS = [700, 1108, 2128, 3000]


def test_autotune_table_and_parity(engine, tmp_path):
    path = str(tmp_path / "pgap.tune")
    table = engine.autotune(WIDTHS, path)
    assert table == pgap.read_tuning(path)
    seen = {(r["test"], r["bits"]) for r in table}
    assert seen == {(k, w) for k in (pgap.FERMAT2, pgap.SPRP, pgap.LUCAS) for w in WIDTHS}
    for r in table:
        assert r["lanes"] in (1, 4, 8, 16, 64) and r["tests_per_s"] > 0 and r["cap"] >= 0
        if r["test"] == pgap.LUCAS:
            assert r["lanes"] in (8, 16, 64)
        # the chosen variant covers the width (strong PRP tuned with base 3)
        limit = r["radix"] * r["digits"] - {pgap.FERMAT2: 4, pgap.SPRP: 8, pgap.LUCAS: 54}[r["test"]]
        assert limit >= r["bits"], r

    tuned = pgap.Engine(0, tuning_path=path)
    try:
        rng = random.Random(7)
        for bits in WIDTHS:
            vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(200)]
            x = vals[0]
            while not ol.probab_prime(x):
                x += 2
            vals.append(x)
            arr = pgap.to_limbs(vals, pgap.limbs_needed(max(vals)))
            for kind, bases in ((pgap.FERMAT2, ()), (pgap.SPRP, (2, 3)), (pgap.LUCAS, ())):
                got = tuned.prp(arr, kind, bases)
                want = ol.verdicts(arr, kind, bases)
                assert np.array_equal(got, want), (bits, kind)
                assert got[-1] == 1
                info = tuned.last_launch()
                want_row = [r for r in table if r["test"] == kind and r["bits"] == bits][0]
                assert (info.radix_bits, info.lanes_per_number) == (want_row["radix"], want_row["lanes"])
    finally:
        tuned.close()


def test_tuning_path_missing_is_an_error(tmp_path):
    with pytest.raises(pgap.PgapError):
        pgap.Engine(0, tuning_path=str(tmp_path / "absent.tune"))
