"""Batch PRP parity: the HIP kernels against the GMP oracle, bit-exact."""
import math
import random

import numpy as np
import pytest

import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu


def special_values():
    v = [0, 1, 2, 3, 4, 5, 7, 9, 15, 341, 561, 645, 1105, 1387, 1729, 1905, 2047, 2465, 2701, 2821, 3277,
         4033, 4369, 4371, 4681, 5461, 6601, 7957, 8321, 25326001, 3215031751, (1 << 31) - 1, (1 << 32) - 5,
         (1 << 61) - 1, (1 << 64) - 59, (1 << 89) - 1, (1 << 127) - 1]
    v += [(1 << p) - 1 for p in (11, 23, 29, 37, 41, 67, 521, 607, 1109, 1279)]
    v += [(1 << (1 << n)) + 1 for n in (5, 6, 7, 8, 9, 10)]
    return v


def candidates(P, D, mlo, mhi, xr, count, seed):
    rng = random.Random(seed)
    K = ol.K(P, D)
    out = []
    while len(out) < count:
        m = rng.randrange(mlo, mhi)
        if math.gcd(m, D) != 1:
            continue
        out.append(m * K + rng.randrange(-xr, xr + 1) | 1)
    return out


def check(engine, vals, kind=pgap.FERMAT2, bases=()):
    limbs = max(pgap.limbs_needed(v) for v in vals)
    arr = pgap.to_limbs(vals, limbs)
    got = engine.prp(arr, kind, bases)
    want = ol.verdicts(arr, kind, bases)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(hex(vals[i]), int(got[i]), int(want[i])) for i in bad[:5]]
    return got


@pytest.mark.parametrize("digits", [40, 78])
def test_sprp_headroom_by_base(digits):
    # Strong PRP multiplies by the base as a scalar: bases up to 8 run with
    # N < R/2^8, larger ones need N < R/2^14.  Numbers of exactly 28*digits - 8
    # bits (largest digits the lazy bound admits among them) run on `digits`
    # digits with bases (3, 5, 7) and on a wider variant with base 61; both
    # match the oracle.
    bits = 28 * digits - 8
    rng = random.Random(digits)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(126)]
    vals += [(1 << bits) - 1, (1 << bits) - 3]
    x = (1 << bits) - 1
    while not ol.probab_prime(x):
        x -= 2
    vals.append(x)
    with pgap.Engine(0) as e:   # untuned: the smallest covering variant
        got = check(e, vals, pgap.SPRP, (3, 5, 7))
        assert got[-1] == 1
        assert e.last_launch().digits == digits
        got = check(e, vals, pgap.SPRP, (3, 61))
        assert got[-1] == 1
        assert e.last_launch().digits > digits


def test_special_values_fermat(engine):
    vals = special_values()
    for v in vals:
        got = engine.prp_ints([v])[0]
        assert got == ol.fermat2(v), hex(v)
    check(engine, vals)


def test_special_values_sprp(engine):
    vals = special_values()
    for bases in ((2,), (3,), (2, 3, 5), (2, 3, 5, 7)):
        check(engine, [v for v in vals if v < (1 << 1700)], pgap.SPRP, bases)


@pytest.mark.parametrize("bits", [30, 60, 64, 100, 200, 400, 700, 706, 900, 1108, 1120, 1300, 1336, 1400,
                                  1650, 1730, 1750, 1900, 2000, 2129, 2180, 2226])
def test_random_sizes(engine, bits):
    rng = random.Random(bits)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(300)]
    p = []
    for v in vals[:6]:
        x = v
        while not ol.probab_prime(x):
            x += 2
        p.append(x)
    got = check(engine, vals + p)
    assert got[-6:].all()


def test_config_families(engine):
    # 503#/210 (N ~ 700 bits) and 809#/1841070 (N ~ 1108 bits), the two
    # sizes the single-lane variants serve.
    a = candidates(503, 210, 1, 2000, 4000, 3000, 11)
    b = candidates(809, 1841070, 10_000_000, 12_000_000, 15000, 3000, 12)
    ga = check(engine, a)
    gb = check(engine, b)
    assert ga.sum() > 0 and gb.sum() > 0
    check(engine, a[:500], pgap.SPRP, (2,))
    check(engine, b[:500], pgap.SPRP, (3, 5))


def test_large_family_and_sprp(engine):
    # 1511#/2190 (N ~ 2109-2129 bits): the wide one-lane variants
    c = candidates(1511, 2190, 1, 1_000_000, 20000, 600, 13)
    assert max(v.bit_length() for v in c) > 2100
    assert check(engine, c).sum() >= 0
    check(engine, c[:200], pgap.SPRP, (3, 5))
    check(engine, c[:100], pgap.SPRP, (2, 61))
    m2203 = (1 << 2203) - 1
    assert engine.prp_ints([m2203])[0] == 1


@pytest.mark.parametrize("bits", [2300, 2490, 2600, 3000, 3320, 3500, 4100, 4250])
def test_quad_variants_random_sizes(engine, bits):
    # wider than one lane holds: four lanes per number up to 3356 bits (radix
    # 2^28, 120 digits), eight to 3580 (128 digits), four again to 4252 (152
    # digits: 4 x 38 ties with 8 x 19 and the fewer lanes win), eight above
    rng = random.Random(bits)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(96)]
    p = []
    for v in vals[:2]:
        x = v
        while not ol.probab_prime(x):
            x += 2
        p.append(x)
    got = check(engine, vals + p)
    assert got[-2:].all()
    quad = bits <= 28 * 120 - 4 or 28 * 128 - 4 < bits <= 28 * 152 - 4
    assert engine.last_launch().lanes_per_number == (4 if quad else 8)


@pytest.mark.parametrize("bits", [3400, 4000, 4240])
def test_quad_38_digit_lanes(bits):
    # the widest quad variant (4 lanes x 38 digits, two wavefronts per SIMD,
    # its own row schedule) against GMP, Fermat and strong PRP, with primes
    # planted among random odd numbers
    rng = random.Random(bits + 38)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(120)]
    for v in vals[:2]:
        x = v
        while not ol.probab_prime(x):
            x += 2
        vals.append(x)
    with pgap.Engine(0, lanes=4) as e:
        got = check(e, vals)
        assert got[-2:].all()
        assert (e.last_launch().lanes_per_number, e.last_launch().digits) == (4, 152)
        check(e, vals[:40] + vals[-2:], pgap.SPRP, (2, 3))
        assert (e.last_launch().lanes_per_number, e.last_launch().digits) == (4, 152)


def test_quad_38_digit_lanes_at_their_bounds():
    # the widest numbers the 4 x 38 variant takes: 28 * 152 - 4 = 4252 bits
    # for Fermat (the squaring's lazy headroom), 4248 for strong PRP, all
    # bits set at the top and a planted prime at each bound
    rng = random.Random(4252)
    with pgap.Engine(0, lanes=4) as e:
        for bits, kind, bases in ((4252, pgap.FERMAT2, ()), (4248, pgap.SPRP, (2,))):
            top = ((1 << bits) - 1) ^ ((1 << (bits - 40)) - 1)   # 40 top bits set
            vals = [top | rng.getrandbits(bits - 40) | 1 for _ in range(48)]
            x = vals[0]
            while not ol.probab_prime(x):
                x -= 2
            vals.append(x)
            assert all(v.bit_length() == bits for v in vals)
            got = check(e, vals, kind, bases)
            assert got[-1] == 1
            assert (e.last_launch().lanes_per_number, e.last_launch().digits) == (4, 152)


@pytest.fixture(scope="module")
def engine8():
    e = pgap.Engine(0, lanes=8)   # force the 8-lanes-per-number variants
    yield e
    e.close()


@pytest.mark.parametrize("bits", [1500, 2128, 2280, 2300, 3000, 3320, 3500, 4100, 4250, 4300, 4360])
def test_octet_variants_random_sizes(engine8, bits):
    # eight lanes per number (radix 2^28, up to 20 digits per lane): up to 4476 bits
    rng = random.Random(bits + 8)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(64)]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    got = check(engine8, vals + [x])
    assert got[-1] == 1
    assert engine8.last_launch().lanes_per_number == 8
    if bits <= 28 * 160 - 14:
        check(engine8, vals[:24] + [x], pgap.SPRP, (3, 5))
        assert engine8.last_launch().lanes_per_number == 8


@pytest.mark.parametrize("bits", [700, 2128, 3000, 4250, 4470])
def test_row_group_variants(bits):
    # sixteen lanes (a whole DPP row) per number, 10 digits per lane: up to
    # 4476 bits Fermat, 4472 strong PRP with bases up to 8
    rng = random.Random(bits + 16)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(64)]
    vals += [(1 << 2281) - 1, (1 << 1279) - 1, (1 << 4096) + 1, 341, 561]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    vals.append(x)
    with pgap.Engine(0, lanes=16) as e:
        got = check(e, vals)
        assert got[-1] == 1 and list(got[64:67]) == [1, 1, 1]
        assert e.last_launch().lanes_per_number == 16
        got = check(e, vals[:24] + vals[64:], pgap.SPRP, (3, 5))
        assert e.last_launch().lanes_per_number == 16
        c = candidates(3001, 2190, 1, 100_000, 40000, 32, 19)   # 4231-4248-bit family
        check(e, c)
        check(e, c[:8], pgap.SPRP, (2, 3))


@pytest.mark.parametrize("bits", [1000, 2128, 2230])
def test_row_group_variants_80_digits(bits):
    # sixteen lanes, 5 digits per lane (80 digits): the scan's variant for
    # 2128-bit numbers while very few m's are in flight
    rng = random.Random(bits + 160)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(64)]
    vals += [(1 << 2203) - 1, (1 << 1279) - 1, 341, 561]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    vals.append(x)
    with pgap.Engine(0, lanes=16) as e:
        got = check(e, vals)
        assert got[-1] == 1 and list(got[64:66]) == [1, 1]
        assert (e.last_launch().lanes_per_number, e.last_launch().digits) == (16, 80)
        check(e, vals[:24] + vals[64:], pgap.SPRP, (3, 5))
        assert (e.last_launch().lanes_per_number, e.last_launch().digits) == (16, 80)


def test_octet_variants_special(engine8):
    vals = [(1 << 2281) - 1, (1 << 3217) - 1, (1 << 4253) - 1, (1 << 4096) + 1, (1 << 4253) + 2, (1 << 4363) - 1]
    got = check(engine8, vals)
    assert list(got[:4]) == [1, 1, 1, 1]
    check(engine8, vals[:4], pgap.SPRP, (2,))
    c = candidates(3001, 2190, 1, 100_000, 40000, 64, 18)   # 4231-4248-bit family
    check(engine8, c)
    check(engine8, c[:16], pgap.SPRP, (3, 5))


def test_quad_variants_special_and_sprp(engine):
    vals = [(1 << 2281) - 1, (1 << 3217) - 1, (1 << 4253) - 1, (1 << 4096) + 1, (1 << 2203) * 3 + 1,
            (1 << 3001) - 1]
    got = check(engine, vals)
    assert list(got[:4]) == [1, 1, 1, 1]
    strong = [v for v in vals if v.bit_length() <= 28 * 160 - 14]  # widest strong-PRP variant
    check(engine, strong, pgap.SPRP, (2,))
    check(engine, strong, pgap.SPRP, (3, 5))
    c = candidates(3001, 2190, 1, 100_000, 40000, 64, 17)   # 4231-4248-bit family
    check(engine, c)
    check(engine, c[:16], pgap.SPRP, (3, 5))


@pytest.mark.parametrize("bits", [100, 700, 1108, 2000])
def test_quad_variants_forced_small(bits):
    # the quad kernels checked where the one-lane kernels also work
    rng = random.Random(bits + 7)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(200)] + [341, 561, 2047, (1 << 521) - 1]
    with pgap.Engine(0, lanes=4) as e:
        check(e, vals)
        assert e.last_launch().lanes_per_number == 4
        check(e, vals[:60], pgap.SPRP, (2, 3))


def test_rejects_what_no_variant_covers(engine):
    with pytest.raises(pgap.PgapError) as ei:
        engine.prp_ints([(1 << 9000) + 1])
    assert ei.value.status == pgap.PGAP_E_TOO_LARGE
    with pytest.raises(pgap.PgapError) as ei:
        engine.prp_ints([101], pgap.SPRP, (64,))
    assert ei.value.status == pgap.PGAP_E_ARG


def test_mixed_limb_padding(engine):
    # the same numbers padded to a wider row give the same verdicts
    vals = candidates(503, 210, 1, 2000, 4000, 500, 5)
    arr = pgap.to_limbs(vals, 40)
    assert (engine.prp(arr) == ol.verdicts(arr)).all()


def test_launch_info(engine):
    vals = candidates(809, 1841070, 10_000_000, 12_000_000, 15000, 1024, 3)
    engine.prp_ints(vals)
    info = engine.last_launch()
    assert info.tests == 1024 and info.kernel_ms > 0 and info.radix_bits == 28
