"""Device memory for Python callers without torch.

``DeviceArray`` holds a hipMalloc'd buffer and copies numpy arrays to and
from it through the HIP runtime that libnebula_mpm.so loaded (the loader
returns the mapped copy for the soname libamdhip64.so.7, so no second runtime
enters the process).  Copies are synchronous.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import mpm

_hip = None
_H2D, _D2H = 1, 2


def runtime():
    global _hip
    if _hip is None:
        mpm.lib()  # the engine library first: its runtime is the one to bind
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _hip = h
    return _hip


def check(rc, what):
    if rc != 0:
        raise mpm.NebulaError(f"{what} failed: {runtime().hipGetErrorString(rc).decode()}")


def synchronize():
    check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device buffer of ``shape`` and ``dtype`` (uninitialised)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        if a.nbytes:
            check(runtime().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy")

    def download(self, out=None):
        out = np.empty(self.shape, self.dtype) if out is None else out
        if out.nbytes != self.nbytes or not out.flags.c_contiguous:
            raise ValueError("output must be a contiguous array of the buffer's size")
        if self.nbytes:
            check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        if _hip is not None:
            self.free()
