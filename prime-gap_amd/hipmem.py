"""Device buffers for Python callers, through the HIP runtime libpgap.so links.

libpgap.so is built against ROCm's libamdhip64.so.7; PyTorch wheels bundle
another copy with the same soname, and one process cannot use both.  Tests
and bench.py therefore allocate device memory here, on the runtime the engine
itself uses, and synchronise with hipDeviceSynchronize.
"""
from __future__ import annotations

import ctypes

import numpy as np

import pgap

_hip = None


def hip():
    global _hip
    if _hip is None:
        pgap.load()  # brings in the runtime libpgap.so was linked against
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        lib.hipGetErrorString.restype = ctypes.c_char_p
        lib.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = lib
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _ck(st: int, what: str):
    if st != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(st).decode()}")


def synchronize():
    _ck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation with a numpy shape/dtype."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ck(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ck(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def copy_from(self, other: "DeviceArray"):
        n = min(self.nbytes, other.nbytes)
        _ck(hip().hipMemcpy(self.ptr, other.ptr, n, D2D), "hipMemcpy D2D")
        return self

    def zero(self):
        _ck(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def to_host(self, rows: int | None = None) -> np.ndarray:
        shape = self.shape if rows is None else (rows,) + self.shape[1:]
        out = np.empty(shape, dtype=self.dtype)
        _ck(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
