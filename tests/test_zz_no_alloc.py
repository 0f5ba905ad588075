"""The allocation-free batch entry, pgap_prp_batch_device_noalloc, against
the allocating one, pgap_prp_batch_device: identical verdicts on two calls
in a row through the C ABI, the second with a smaller input (its survivor
scratch keeps the first call's contents), and the oracle's verdicts."""
import random

import numpy as np
import pytest

import hipmem
import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu


def numbers(count, bits, seed):
    rng = random.Random(seed)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(count - 8)]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    # primes, a strong base-2 pseudoprime (a composite Mersenne number) and 2
    vals += [x, x, (1 << 1109) - 1 if bits > 1109 else 2047, 2, 341, 561, 3, 25326001]
    return vals


@pytest.mark.parametrize("kind,bases", [(pgap.FERMAT2, ()), (pgap.SPRP, (2, 3, 5)), (pgap.LUCAS, ()),
                                        (pgap.BPSW, ())])
def test_noalloc_matches_allocating_entry(engine, kind, bases):
    for count, bits, seed in ((6000, 1150, 1), (700, 2128, 2)):   # the second call is smaller
        vals = numbers(count, bits, seed)
        limbs = pgap.limbs_needed(max(vals))
        arr = pgap.to_limbs(vals, limbs)
        d_n = hipmem.DeviceArray.from_host(arr)
        d_a = hipmem.DeviceArray(count, np.uint8)
        d_b = hipmem.DeviceArray(count, np.uint8)
        max_bits = max(v.bit_length() for v in vals)
        engine.prp_device(d_n.ptr, limbs, count, d_a.ptr, kind, bases, max_bits)
        engine.sync()
        st = engine.prp_device_noalloc(d_n.ptr, limbs, count, d_b.ptr, kind, bases, max_bits)
        assert st == pgap.PGAP_OK, engine.last_error()
        engine.sync()
        a, b = d_a.to_host(), d_b.to_host()
        assert np.array_equal(a, b), (kind, count)
        assert np.array_equal(b, ol.verdicts(arr, kind, bases)), (kind, count)
        for d in (d_n, d_a, d_b):
            d.free()


def test_noalloc_refuses_what_its_scratch_cannot_hold(engine):
    d_n = hipmem.DeviceArray((4, 2), np.uint32)
    d_o = hipmem.DeviceArray(4, np.uint8)
    # nothing is launched for either call
    assert engine.prp_device_noalloc(d_n.ptr, 2, pgap.NOALLOC_MAX_COUNT + 1, d_o.ptr, pgap.BPSW, (), 64) == \
        pgap.PGAP_E_TOO_LARGE
    assert "exceed" in engine.last_error()
    assert engine.prp_device_noalloc(d_n.ptr, 2, 4, d_o.ptr, pgap.BPSW, (), 0) == pgap.PGAP_E_ARG
    d_n.free()
    d_o.free()
