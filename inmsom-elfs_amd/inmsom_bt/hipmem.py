"""Device arrays on the HIP runtime libinmsom_bt.so itself links
(/opt/rocm's libamdhip64), for the entry points that take device
addresses.  A second HIP runtime in the same process (the one PyTorch
bundles) does not see the GPU once this one holds it, so device memory for
those entry points comes from here."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _capi

_H2D, _D2H = 1, 2
_hip = None


def hip():
    """libamdhip64 as already loaded by libinmsom_bt.so."""
    global _hip
    if _hip is None:
        _capi.lib()
        L = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipSetDevice.argtypes = [C.c_int]
        L.hipGetErrorString.argtypes = [C.c_int]
        L.hipGetErrorString.restype = C.c_char_p
        for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipSetDevice, L.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = L
    return _hip


def _ok(code: int, what: str):
    if code != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(code).decode()}")


class DeviceArray:
    """A C-contiguous float64 array in device memory on one GPU."""

    dtype = np.dtype(np.float64)

    def __init__(self, shape: Sequence[int], device: int = 0):
        self.shape = tuple(int(s) for s in shape)
        self.device = int(device)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _ok(hip().hipSetDevice(self.device), "hipSetDevice")
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape, device)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _ok(hip().hipSetDevice(self.device), "hipSetDevice")
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _ok(hip().hipSetDevice(self.device), "hipSetDevice")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _ok(hip().hipSetDevice(self.device), "hipSetDevice")
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
