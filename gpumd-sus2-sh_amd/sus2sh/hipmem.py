"""Minimal device-memory plumbing through the HIP runtime that libsus2sh.so itself uses.

`DeviceArray` wraps hipMalloc'd memory with numpy-style upload/download so callers that do
not use torch can drive the C ABI.  Only memory management lives here; every computation
is done by the engine's kernels.
"""
import ctypes

import numpy as np

from ._capi import lib

_hip = None


def hip():
    global _hip
    if _hip is not None:
        return _hip
    lib()  # the engine library pulls in the HIP runtime it was linked with
    H = ctypes.CDLL("libamdhip64.so.7")
    P = ctypes.POINTER
    H.hipMalloc.argtypes = [P(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMalloc.restype = ctypes.c_int
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipFree.restype = ctypes.c_int
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipMemcpy.restype = ctypes.c_int
    H.hipDeviceSynchronize.argtypes = []
    H.hipDeviceSynchronize.restype = ctypes.c_int
    H.hipGetDeviceCount.argtypes = [P(ctypes.c_int)]
    H.hipGetDeviceCount.restype = ctypes.c_int
    H.hipEventCreate.argtypes = [P(ctypes.c_void_p)]
    H.hipEventCreate.restype = ctypes.c_int
    H.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventRecord.restype = ctypes.c_int
    H.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    H.hipEventSynchronize.restype = ctypes.c_int
    H.hipEventElapsedTime.argtypes = [P(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventElapsedTime.restype = ctypes.c_int
    _hip = H
    return H


H2D, D2H = 1, 2


def _ok(code, what):
    if code != 0:
        raise RuntimeError(f"{what} failed with hipError {code}")


def device_count():
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """Contiguous device buffer with a fixed dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        d.upload(np.zeros(d.shape, d.dtype))
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D),
                "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, D2H),
                "hipMemcpy D2H")
        return out

    def data_ptr(self):
        return self.ptr.value

    def numel(self):
        return int(np.prod(self.shape))

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
