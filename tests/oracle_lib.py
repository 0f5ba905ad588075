"""ctypes wrapper of oracle/liboracle.so -- the GMP checker.  Tests only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_SO)
    vp, u32, u64, i64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_size_t
    lib.or_fermat2.argtypes = [vp, ctypes.c_int]
    lib.or_sprp.argtypes = [vp, ctypes.c_int, u32]
    lib.or_probab_prime.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.or_strong_lucas.argtypes = [vp, ctypes.c_int]
    lib.or_bpsw.argtypes = [vp, ctypes.c_int]
    lib.or_verdicts.argtypes = [vp, sz, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int]
    lib.or_verdicts.restype = None
    lib.or_K.argtypes = [u32, u32, vp, ctypes.c_int]
    lib.or_K_log.argtypes = [u32, u32]
    lib.or_K_log.restype = ctypes.c_double
    lib.or_prev_prime_dist.argtypes = [u32, u32, u64, i64]
    lib.or_prev_prime_dist.restype = i64
    lib.or_next_prime_dist.argtypes = [u32, u32, u64, i64]
    lib.or_next_prime_dist.restype = i64
    lib.or_test_interval.argtypes = [u32, u32, u64, u32, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp]
    lib.or_test_interval.restype = None
    lib.or_sieve_new.argtypes = [u32, u32, u32, u64]
    lib.or_sieve_new.restype = vp
    lib.or_sieve_free.argtypes = [vp]
    lib.or_sieve_free.restype = None
    lib.or_sieve_m.argtypes = [vp, u64, vp, vp, vp, vp, ctypes.c_int]
    lib.or_sieve_new_range.argtypes = [u32, u32, u32, u64, u64]
    lib.or_sieve_new_range.restype = vp
    lib.or_sieve_removed.argtypes = [vp, u64, u64, vp, vp, ctypes.c_int]
    lib.or_sieve_removed.restype = ctypes.c_int
    lib.or_bench_scan.argtypes = [u32, u32, u32, vp, sz, vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    lib.or_bench_scan.restype = sz
    _lib = lib
    return lib


def _limbs(x: int) -> np.ndarray:
    n = max(1, (x.bit_length() + 31) // 32)
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype="<u4").copy()


def fermat2(x: int) -> int:
    a = _limbs(x)
    return load().or_fermat2(a.ctypes.data, len(a))


def sprp(x: int, base: int) -> int:
    a = _limbs(x)
    return load().or_sprp(a.ctypes.data, len(a), base)


def probab_prime(x: int, reps: int = 25) -> int:
    a = _limbs(x)
    return load().or_probab_prime(a.ctypes.data, len(a), reps)


def strong_lucas(x: int) -> int:
    a = _limbs(x)
    return load().or_strong_lucas(a.ctypes.data, len(a))


def bpsw(x: int) -> int:
    a = _limbs(x)
    return load().or_bpsw(a.ctypes.data, len(a))


def verdicts(n: np.ndarray, kind: int = 0, bases=(), threads: int = 8) -> np.ndarray:
    n = np.ascontiguousarray(n, dtype=np.uint32)
    out = np.zeros(n.shape[0], dtype=np.uint8)
    b = np.ascontiguousarray(bases, dtype=np.uint32)
    load().or_verdicts(n.ctypes.data, n.shape[0], n.shape[1], kind, b.ctypes.data if len(b) else None, len(b),
                       out.ctypes.data, threads)
    return out


def K(P: int, D: int) -> int:
    buf = np.zeros(4096, dtype=np.uint32)
    n = load().or_K(P, D, buf.ctypes.data, len(buf))
    if n < 0:
        raise ValueError("d does not divide P#")
    return int.from_bytes(buf[:n].tobytes(), "little")


def K_log(P: int, D: int) -> float:
    return load().or_K_log(P, D)


def prev_prime_dist(P, D, m, start=0) -> int:
    return load().or_prev_prime_dist(P, D, m, start)


def next_prime_dist(P, D, m, start=0) -> int:
    return load().or_next_prime_dist(P, D, m, start)


def test_interval(P, D, m, SL, prev, nxt, kind=0) -> dict:
    p = np.ascontiguousarray(prev, dtype=np.int32)
    q = np.ascontiguousarray(nxt, dtype=np.int32)
    out = np.zeros(7, dtype=np.int64)
    load().or_test_interval(P, D, m, SL, p.ctypes.data, len(p), q.ctypes.data, len(q), kind, out.ctypes.data)
    keys = ("prev_p", "next_p", "tests", "prev_fallback", "next_fallback", "p_tests", "n_tests")
    return dict(zip(keys, (int(v) for v in out)))


def scan_many(P, D, SL, ms, lists, kind=0, threads=8) -> np.ndarray:
    """gap_test_simple over many m's on `threads` threads -> rows [n][3]
    (prev_p, next_p, prp tests)."""
    n = len(ms)
    m = np.ascontiguousarray(ms, dtype=np.uint64)
    npv = np.array([len(p) for p, _ in lists], dtype=np.uint32)
    nnx = np.array([len(q) for _, q in lists], dtype=np.uint32)
    poff = np.zeros(n, dtype=np.uint64)
    qoff = np.zeros(n, dtype=np.uint64)
    poff[1:] = np.cumsum(npv[:-1], dtype=np.uint64)
    qoff[1:] = np.cumsum(nnx[:-1], dtype=np.uint64)
    pv = np.concatenate([np.asarray(p, np.int32) for p, _ in lists] + [np.zeros(1, np.int32)])
    nx = np.concatenate([np.asarray(q, np.int32) for _, q in lists] + [np.zeros(1, np.int32)])
    rows = np.zeros((n, 3), dtype=np.int64)
    load().or_bench_scan(P, D, SL, m.ctypes.data, n, pv.ctypes.data, poff.ctypes.data, npv.ctypes.data,
                         nx.ctypes.data, qoff.ctypes.data, nnx.ctypes.data, kind, threads, rows.ctypes.data)
    return rows


class Sieve:
    """Unknowns of m*K+x, x in [-SL, SL], free of primes <= max_prime."""

    def __init__(self, P, D, SL, max_prime, above: int = 0):
        """above > 0: keep only the primes in (above, max_prime] (for removed())."""
        self.lib = load()
        self.h = self.lib.or_sieve_new_range(P, D, SL, above, max_prime)
        if not self.h:
            raise ValueError("bad sieve parameters")
        self.SL = SL

    def __call__(self, m: int):
        cap = self.SL + 1
        prev = np.zeros(cap, dtype=np.int32)
        nxt = np.zeros(cap, dtype=np.int32)
        a = ctypes.c_int()
        b = ctypes.c_int()
        r = self.lib.or_sieve_m(self.h, m, prev.ctypes.data, ctypes.byref(a), nxt.ctypes.data, ctypes.byref(b), cap)
        assert r == 0
        return prev[: a.value].copy(), nxt[: b.value].copy()

    def removed(self, m: int, lo: int):
        """[(x, p)]: offsets some prime p in (lo, max_prime] removes, p the first."""
        cap = 2 * self.SL + 1
        xs = np.zeros(cap, dtype=np.int64)
        ps = np.zeros(cap, dtype=np.uint64)
        n = self.lib.or_sieve_removed(self.h, m, lo, xs.ctypes.data, ps.ctypes.data, cap)
        assert n >= 0
        return [(int(xs[i]), int(ps[i])) for i in range(n)]

    def close(self):
        if self.h:
            self.lib.or_sieve_free(self.h)
            self.h = None

    def __del__(self):
        self.close()
