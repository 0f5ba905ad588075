"""ctypes binding of libpgap.so (the C ABI declared in include/pgap.h).

This is the binding a Python caller of the reference would add (the
reference's own Python tester, gap_test.py, calls gmpy2 per candidate,
gap_utils_primes.py:33-47); tests and bench.py use it to drive the HIP engine.
There is no CPU fallback: if libpgap.so is missing or no GPU is present the
calls raise PgapError.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgap.so")

PGAP_OK = 0
PGAP_E_ARG = -1
PGAP_E_DEVICE = -2
PGAP_E_OOM = -3
PGAP_E_TOO_LARGE = -4
PGAP_E_STATE = -5
PGAP_E_TIMEOUT = -6

FERMAT2 = 0
SPRP = 1
LUCAS = 2   # strong Lucas, Selfridge's method A
BPSW = 3    # strong base 2 + strong Lucas: mpz_probab_prime_p's verdict

CONFIRM_NONE = 0   # a side's prime is its first Fermat base-2 PRP
CONFIRM_SPRP2 = 1  # ... confirmed by a strong base-2 test (the reference GPU tester's verdict)
CONFIRM_BPSW = 2   # ... confirmed by Baillie-PSW (gap_test_simple's verdict; the default)

ROW_PREV_FALLBACK = 1
ROW_NEXT_FALLBACK = 2
ROW_NEXT_SKIPPED = 4
ROW_PREV_SKIPPED = 8
SIDES_BOTH, SIDES_PREV, SIDES_NEXT = 0, 1, 2


def read_tuning(path: str) -> List[dict]:
    """Parse an autotuner table (the text format pgap_autotune writes)."""
    keys = ("test", "bits", "radix", "digits", "lanes", "cap", "tests_per_s")
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            rows.append({k: (float(v) if k == "tests_per_s" else int(v)) for k, v in zip(keys, parts)})
    return rows


# Every symbol include/pgap.h declares; tests check the library exports them.
EXPORTED = (
    "pgap_create", "pgap_destroy", "pgap_last_error", "pgap_version",
    "pgap_prp_batch", "pgap_prp_batch_device", "pgap_prp_batch_mixed", "pgap_sync",
    "pgap_last_launch", "pgap_family", "pgap_scan_submit", "pgap_scan_poll",
    "pgap_scan_in_flight", "pgap_scan_config", "pgap_build_candidates_device", "pgap_filter_candidates",
    "pgap_sieve", "pgap_scan_stats", "pgap_autotune", "pgap_scan_side_skip",
    "pgap_prp_batch_device_noalloc", "pgap_mad_peak",
)
NOALLOC_MAX_COUNT = 1 << 20


class PgapError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"pgap status {status}: {msg}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("hbm_budget", ctypes.c_size_t),
        ("tuning_path", ctypes.c_char_p),
        ("radix", ctypes.c_int),
        ("lanes_per_number", ctypes.c_int),
        ("waves_per_simd", ctypes.c_int),
    ]


class LaunchInfo(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("digits", ctypes.c_uint32),
        ("radix_bits", ctypes.c_uint32),
        ("lanes_per_number", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
        ("tests", ctypes.c_uint64),
        ("graph", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class ScanReq(ctypes.Structure):
    _fields_ = [
        ("family", ctypes.c_int),
        ("m", ctypes.c_uint64),
        ("prev", ctypes.POINTER(ctypes.c_int32)),
        ("n_prev", ctypes.c_uint32),
        ("next", ctypes.POINTER(ctypes.c_int32)),
        ("n_next", ctypes.c_uint32),
        ("tag", ctypes.c_uint64),
        ("sides", ctypes.c_uint32),
    ]


class ScanRow(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_uint64),
        ("tag", ctypes.c_uint64),
        ("prev_p", ctypes.c_int32),
        ("next_p", ctypes.c_int32),
        ("p_tests", ctypes.c_uint32),
        ("n_tests", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("family", ctypes.c_uint32),
        ("t_submit", ctypes.c_double),
        ("t_done", ctypes.c_double),
    ]


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libpgap.so (once).  Raises PgapError if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PgapError(PGAP_E_DEVICE, f"{path} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    lib.pgap_create.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(vp)]
    lib.pgap_destroy.argtypes = [vp]
    lib.pgap_destroy.restype = None
    lib.pgap_last_error.argtypes = [vp]
    lib.pgap_last_error.restype = ctypes.c_char_p
    lib.pgap_version.restype = ctypes.c_char_p
    lib.pgap_prp_batch.argtypes = [vp, u32, vp, sz, ctypes.c_int, vp, u32, vp]
    lib.pgap_prp_batch_device.argtypes = [vp, u32, vp, sz, ctypes.c_int, vp, u32, vp, u32]
    lib.pgap_prp_batch_device_noalloc.argtypes = [vp, u32, vp, sz, ctypes.c_int, vp, u32, vp, u32]
    lib.pgap_prp_batch_mixed.argtypes = [vp, vp, vp, vp, sz, ctypes.c_int, vp, u32, vp]
    lib.pgap_sync.argtypes = [vp]
    lib.pgap_last_launch.argtypes = [vp, ctypes.POINTER(LaunchInfo)]
    lib.pgap_family.argtypes = [vp, u32, u32, u32, ctypes.POINTER(ctypes.c_int)]
    lib.pgap_scan_submit.argtypes = [vp, ctypes.POINTER(ScanReq)]
    lib.pgap_scan_poll.argtypes = [vp, ctypes.POINTER(ScanRow), sz, ctypes.POINTER(sz), ctypes.c_int]
    lib.pgap_scan_in_flight.argtypes = [vp]
    lib.pgap_scan_in_flight.restype = sz
    lib.pgap_scan_config.argtypes = [vp, u32, u32]
    lib.pgap_scan_side_skip.argtypes = [vp, ctypes.c_int, ctypes.c_float]
    lib.pgap_build_candidates_device.argtypes = [vp, ctypes.c_int, vp, vp, sz, u32, vp]
    lib.pgap_filter_candidates.argtypes = [vp, ctypes.c_int, vp, vp, sz, u32, vp]
    lib.pgap_sieve.argtypes = [vp, ctypes.c_int, vp, sz, u64, vp, vp, vp, vp]
    lib.pgap_scan_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u32)]
    lib.pgap_autotune.argtypes = [vp, ctypes.POINTER(u32), ctypes.c_size_t, ctypes.c_char_p]
    lib.pgap_mad_peak.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(u64)]
    for name in EXPORTED:
        f = getattr(lib, name)
        if name not in ("pgap_destroy", "pgap_last_error", "pgap_version", "pgap_scan_in_flight"):
            f.restype = ctypes.c_int
    _lib = lib
    return lib


# ---------------------------------------------------------------- helpers

def limbs_needed(x: int) -> int:
    return max(1, (x.bit_length() + 31) // 32)


def to_limbs(values: Iterable[int], limbs: int) -> np.ndarray:
    """Row-major [count][limbs] little-endian uint32 words."""
    vals = list(values)
    out = np.zeros((len(vals), limbs), dtype=np.uint32)
    for i, v in enumerate(vals):
        if v < 0 or v.bit_length() > 32 * limbs:
            raise ValueError(f"value {i} does not fit {limbs} limbs")
        b = v.to_bytes(4 * limbs, "little")
        out[i] = np.frombuffer(b, dtype="<u4")
    return out


def from_limbs(row: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(row, dtype="<u4").tobytes(), "little")


class Engine:
    """One pgap_engine.  Mirrors test_runner_t (miller_rabin.h:293-383)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, radix: int = 0, lanes: int = 0,
                 waves_per_simd: int = 0, tuning_path: Optional[str] = None):
        self.lib = load()
        opt = Options(device, hbm_budget, tuning_path.encode() if tuning_path else None, radix, lanes,
                      waves_per_simd)
        h = ctypes.c_void_p()
        st = self.lib.pgap_create(ctypes.byref(opt), ctypes.byref(h))
        self.h = h
        if st != PGAP_OK:
            msg = self.lib.pgap_last_error(h).decode() if h else "pgap_create failed"
            self.close()
            raise PgapError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pgap_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != PGAP_OK:
            raise PgapError(st, self.lib.pgap_last_error(self.h).decode())

    def last_error(self) -> str:
        return self.lib.pgap_last_error(self.h).decode()

    def prp(self, numbers: np.ndarray, kind: int = FERMAT2, bases: Sequence[int] = ()) -> np.ndarray:
        """numbers: [count][limbs] uint32 -> verdicts uint8[count]."""
        n = np.ascontiguousarray(numbers, dtype=np.uint32)
        if n.ndim != 2:
            raise ValueError("numbers must be 2-D [count][limbs]")
        out = np.zeros(n.shape[0], dtype=np.uint8)
        b = np.ascontiguousarray(bases, dtype=np.uint32)
        self._check(self.lib.pgap_prp_batch(self.h, n.shape[1], n.ctypes.data, n.shape[0], kind,
                                            b.ctypes.data if len(b) else None, len(b), out.ctypes.data))
        return out

    def prp_ints(self, values: Sequence[int], kind: int = FERMAT2, bases: Sequence[int] = ()) -> np.ndarray:
        limbs = max(limbs_needed(v) for v in values)
        return self.prp(to_limbs(values, limbs), kind, bases)

    def prp_device(self, d_n: int, limbs: int, count: int, d_out: int, kind: int = FERMAT2,
                   bases: Sequence[int] = (), max_bits: int = 0):
        b = np.ascontiguousarray(bases, dtype=np.uint32)
        self._check(self.lib.pgap_prp_batch_device(self.h, limbs, d_n, count, kind,
                                                   b.ctypes.data if len(b) else None, len(b), d_out, max_bits))

    def prp_device_noalloc(self, d_n: int, limbs: int, count: int, d_out: int, kind: int = FERMAT2,
                           bases: Sequence[int] = (), max_bits: int = 0) -> int:
        """pgap_prp_batch_device_noalloc; returns the status (PGAP_OK or an error)."""
        b = np.ascontiguousarray(bases, dtype=np.uint32)
        return self.lib.pgap_prp_batch_device_noalloc(self.h, limbs, d_n, count, kind,
                                                      b.ctypes.data if len(b) else None, len(b), d_out, max_bits)

    def prp_mixed(self, values: Sequence[np.ndarray], kind: int = FERMAT2, bases: Sequence[int] = ()) -> np.ndarray:
        """values: list of 1-D uint32 limb arrays of varying length."""
        limbs = np.array([len(v) for v in values], dtype=np.uint32)
        offset = np.zeros(len(values) + 1, dtype=np.uint64)
        offset[1:] = np.cumsum(limbs, dtype=np.uint64)
        words = np.concatenate([np.asarray(v, dtype=np.uint32) for v in values]) if values else np.zeros(1, np.uint32)
        out = np.zeros(len(values), dtype=np.uint8)
        b = np.ascontiguousarray(bases, dtype=np.uint32)
        self._check(self.lib.pgap_prp_batch_mixed(self.h, words.ctypes.data, offset.ctypes.data, limbs.ctypes.data,
                                                  len(values), kind, b.ctypes.data if len(b) else None, len(b),
                                                  out.ctypes.data))
        return out

    def prp_mixed_packed(self, words: np.ndarray, offset: np.ndarray, limbs: np.ndarray, kind: int = FERMAT2,
                         bases: Sequence[int] = ()) -> np.ndarray:
        """Number i = words[offset[i] : offset[i] + limbs[i]] (uint32 words)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        o = np.ascontiguousarray(offset, dtype=np.uint64)
        lb = np.ascontiguousarray(limbs, dtype=np.uint32)
        out = np.zeros(len(lb), dtype=np.uint8)
        b = np.ascontiguousarray(bases, dtype=np.uint32)
        self._check(self.lib.pgap_prp_batch_mixed(self.h, w.ctypes.data, o.ctypes.data, lb.ctypes.data, len(lb), kind,
                                                  b.ctypes.data if len(b) else None, len(b), out.ctypes.data))
        return out

    def sync(self):
        self._check(self.lib.pgap_sync(self.h))

    def last_launch(self) -> LaunchInfo:
        info = LaunchInfo()
        self._check(self.lib.pgap_last_launch(self.h, ctypes.byref(info)))
        return info

    # ---- streaming scan
    def family(self, P: int, D: int, sieve_length: int) -> int:
        fid = ctypes.c_int()
        self._check(self.lib.pgap_family(self.h, P, D, sieve_length, ctypes.byref(fid)))
        return fid.value

    def build_candidates_device(self, family: int, m: np.ndarray, x: np.ndarray, limbs: int, d_out: int):
        mm = np.ascontiguousarray(m, dtype=np.uint64)
        xx = np.ascontiguousarray(x, dtype=np.int64)
        assert len(mm) == len(xx)
        self._check(self.lib.pgap_build_candidates_device(self.h, family, mm.ctypes.data, xx.ctypes.data, len(mm),
                                                          limbs, d_out))

    def filter_candidates(self, family: int, m: np.ndarray, x: np.ndarray, limit: int) -> np.ndarray:
        mm = np.ascontiguousarray(m, dtype=np.uint64)
        xx = np.ascontiguousarray(x, dtype=np.int64)
        keep = np.zeros(len(mm), dtype=np.uint8)
        self._check(self.lib.pgap_filter_candidates(self.h, family, mm.ctypes.data, xx.ctypes.data, len(mm), limit,
                                                    keep.ctypes.data))
        return keep

    def sieve(self, family: int, ms, max_prime: int, sieve_length: int):
        """-> list of (prev distances, next distances) per m (device sieve)."""
        mm = np.ascontiguousarray(ms, dtype=np.uint64)
        n = len(mm)
        prev = np.zeros((n, sieve_length), dtype=np.int32)
        nxt = np.zeros((n, sieve_length), dtype=np.int32)
        npv = np.zeros(n, dtype=np.uint32)
        nnx = np.zeros(n, dtype=np.uint32)
        self._check(self.lib.pgap_sieve(self.h, family, mm.ctypes.data, n, max_prime, prev.ctypes.data,
                                        npv.ctypes.data, nxt.ctypes.data, nnx.ctypes.data))
        return [(prev[i, : npv[i]].copy(), nxt[i, : nnx[i]].copy()) for i in range(n)]

    def scan_config(self, depth: int, confirm: int = CONFIRM_BPSW):
        """Speculative depth cap and the confirmation (CONFIRM_*)."""
        self._check(self.lib.pgap_scan_config(self.h, depth, confirm))

    def scan_side_skip(self, enable: bool, min_merit: float = 18.0):
        """gap_test_gpu's prev-first order with side-skip below min_merit/2 - 2."""
        self._check(self.lib.pgap_scan_side_skip(self.h, 1 if enable else 0, float(min_merit)))

    def scan_submit(self, family: int, m: int, prev: np.ndarray, nxt: np.ndarray, tag: int = 0,
                    sides: int = 0) -> int:
        """Returns PGAP_OK, or PGAP_E_OOM when admission is refused; raises
        otherwise.  sides: SIDES_BOTH, SIDES_PREV or SIDES_NEXT."""
        p = np.ascontiguousarray(prev, dtype=np.int32)
        q = np.ascontiguousarray(nxt, dtype=np.int32)
        req = ScanReq(family, m, p.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(p),
                      q.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(q), tag, sides)
        st = self.lib.pgap_scan_submit(self.h, ctypes.byref(req))
        if st not in (PGAP_OK, PGAP_E_OOM):
            self._check(st)
        return st

    def scan_poll(self, cap: int = 4096, wait: bool = True, timeout_ms: int = -1) -> list:
        """Finished rows.  wait: block until a row is ready or nothing is in
        flight, at most timeout_ms (< 0: no limit; an empty list when it
        passes with nothing ready)."""
        rows = (ScanRow * cap)()
        n = ctypes.c_size_t()
        st = self.lib.pgap_scan_poll(self.h, rows, cap, ctypes.byref(n), timeout_ms if wait else 0)
        if st != PGAP_E_TIMEOUT:
            self._check(st)
        return [rows[i] for i in range(n.value)]

    def scan_in_flight(self) -> int:
        return self.lib.pgap_scan_in_flight(self.h)

    def autotune(self, widths: Sequence[int], path: Optional[str] = None) -> List[dict]:
        """Time the kernel families on each width and keep the fastest per
        (test, width); returns the table and writes it to `path` if given
        (load it later with Engine(tuning_path=path))."""
        arr = (ctypes.c_uint32 * len(widths))(*[int(w) for w in widths])
        tmp = None
        if path is None:
            import tempfile
            fd, tmp = tempfile.mkstemp(suffix=".tune")
            os.close(fd)
        try:
            self._check(self.lib.pgap_autotune(self.h, arr, len(widths), (path or tmp).encode()))
            return read_tuning(path or tmp)
        finally:
            if tmp:
                os.unlink(tmp)

    def mad_peak(self, iters: int = 4096) -> dict:
        """Sustained v_mad_u64_u32 rate of the device (the roofline denominator)."""
        rate, ms, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        self._check(self.lib.pgap_mad_peak(self.h, iters, ctypes.byref(rate), ctypes.byref(ms), ctypes.byref(n)))
        return {"mac_per_s": rate.value, "kernel_ms": ms.value, "macs": n.value}

    def scan_stats(self) -> dict:
        t, r, s = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        self._check(self.lib.pgap_scan_stats(self.h, ctypes.byref(t), ctypes.byref(r), ctypes.byref(s)))
        return {"tests_issued": t.value, "rounds": r.value, "slots": s.value}
