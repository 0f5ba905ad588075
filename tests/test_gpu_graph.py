"""Steady-state graph replay: a batch launch repeated with the same arguments
is captured into a hipGraph on its second occurrence and replayed after.
Every replay must give the eager verdicts, with eager work (another engine's
launches, buffer clears, host synchronisation) between the replays."""
import random

import numpy as np
import pytest

import hipmem
import oracle_lib as ol
import pgap

pytestmark = pytest.mark.gpu


def rows(bits, count, seed):
    rng = random.Random(seed)
    vals = [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(count)]
    x = vals[0]
    while not ol.probab_prime(x):
        x += 2
    vals[-1] = x
    return pgap.to_limbs(vals, (bits + 31) // 32)


@pytest.mark.parametrize("kind,bases,bits", [(pgap.FERMAT2, (), 1108), (pgap.SPRP, (3, 5), 1108),
                                             (pgap.SPRP, (2, 3, 5), 2128), (pgap.FERMAT2, (), 4247)])
def test_replay_matches_eager(kind, bases, bits):
    arr = rows(bits, 3000, bits + len(bases))
    want = ol.verdicts(arr, kind, bases)
    d_n = hipmem.DeviceArray.from_host(arr)
    d_o = hipmem.DeviceArray(len(arr), np.uint8)
    other = rows(700, 512, 5)
    o_n = hipmem.DeviceArray.from_host(other)
    o_o = hipmem.DeviceArray(len(other), np.uint8)
    with pgap.Engine(0) as e, pgap.Engine(0) as side:
        graphs = []
        for i in range(6):
            d_o.zero()
            hipmem.synchronize()
            e.prp_device(d_n.ptr, arr.shape[1], len(arr), d_o.ptr, kind, bases, bits)
            e.sync()
            graphs.append(e.last_launch().graph)
            assert np.array_equal(d_o.to_host(), want), (i, graphs)
            # eager work between replays: another engine's launches and a clear
            side.prp_device(o_n.ptr, other.shape[1], len(other), o_o.ptr, pgap.FERMAT2, (), 700)
            side.sync()
            o_o.zero()
        assert graphs == [0, 1, 1, 1, 1, 1]
        # different arguments: eager again, then a new graph
        e.prp_device(d_n.ptr, arr.shape[1], len(arr) - 1, d_o.ptr, kind, bases, bits)
        e.sync()
        assert e.last_launch().graph == 0
