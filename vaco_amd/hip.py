"""Minimal device-memory plumbing over the HIP runtime the library links.

The library links ROCm's ``libamdhip64.so.7``; PyTorch wheels bundle their
own HIP runtime, and two runtimes in one process do not share devices. So
processes that load ``libvaco_av1pf.so`` allocate and copy device memory
through the already-loaded runtime here instead of through torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import av1pf

_rt = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _rt
    if _rt is None:
        av1pf.lib()  # load the library first so the soname resolves to its runtime
        R = C.CDLL("libamdhip64.so.7")
        R.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        R.hipFree.argtypes = [C.c_void_p]
        R.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        R.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        R.hipDeviceSynchronize.argtypes = []
        R.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        R.hipSetDevice.argtypes = [C.c_int]
        R.hipGetErrorString.argtypes = [C.c_int]
        R.hipGetErrorString.restype = C.c_char_p
        R.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        R.hipStreamDestroy.argtypes = [C.c_void_p]
        R.hipStreamSynchronize.argtypes = [C.c_void_p]
        _rt = R
    return _rt


def check(e: int, what: str):
    if e != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(e).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    e = rt().hipGetDeviceCount(C.byref(n))
    return n.value if e == 0 else 0


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation holding a 2-D uint16 plane (or any byte size)."""

    def __init__(self, shape, dtype=np.uint16):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(rt().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(rt().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, H2D), "hipMemcpy H2D")

    def zero(self):
        """Clear the allocation (synchronous copy of zeros from the host)."""
        self.upload(np.zeros(self.shape, self.dtype))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority (caller streams for
    the library's *_device entry points; `handle` is the hipStream_t)."""

    def __init__(self):
        h = C.c_void_p()
        check(rt().hipStreamCreateWithFlags(C.byref(h), 1), "hipStreamCreateWithFlags")
        self.handle = h.value

    def synchronize(self):
        check(rt().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            rt().hipStreamDestroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
