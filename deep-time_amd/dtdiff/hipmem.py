"""Device arrays for the device-resident entry points, through the HIP runtime the library
itself links (libamdhip64.so.7), so one process holds one runtime.  Plumbing only."""
from __future__ import annotations

import ctypes

import numpy as np

from . import load

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        load()  # the library first: its runtime is the one resolved below
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipSetDevice, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """n elements of `dtype` in device memory (freed with the object)."""

    def __init__(self, n: int, dtype=np.float64, device: int = 0):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        _ok(hip().hipSetDevice(device), "hipSetDevice")
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, device)
        d.upload(a)
        return d

    def data_ptr(self) -> int:
        return self.ptr.value

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D),
            "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, D2H),
            "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        assert other.nbytes == self.nbytes
        _ok(hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy D2D")

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
