"""Device buffers, streams and events through the HIP runtime libaccp uses.

Plumbing for tests and the benchmark: the same libamdhip64 instance that
libaccp.so is linked against (found by soname after libaccp is loaded), so a
process never mixes two HIP runtimes.
"""
import ctypes
from ctypes import POINTER, byref, c_float, c_int, c_size_t, c_void_p

import numpy as np

H2D, D2H, D2D = 1, 2, 3

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        from .abi import load

        load()  # libaccp first: its NEEDED libamdhip64.so.7 is then the loaded one
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
        L.hipFree.argtypes = [c_void_p]
        L.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        L.hipMemcpyAsync.argtypes = [c_void_p, c_void_p, c_size_t, c_int, c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [POINTER(c_void_p)]
        L.hipStreamDestroy.argtypes = [c_void_p]
        L.hipStreamSynchronize.argtypes = [c_void_p]
        L.hipEventCreate.argtypes = [POINTER(c_void_p)]
        L.hipEventDestroy.argtypes = [c_void_p]
        L.hipEventRecord.argtypes = [c_void_p, c_void_p]
        L.hipEventSynchronize.argtypes = [c_void_p]
        L.hipEventElapsedTime.argtypes = [POINTER(c_float), c_void_p, c_void_p]
        L.hipGetErrorString.argtypes = [c_int]
        L.hipGetErrorString.restype = ctypes.c_char_p
        _HIP = L
    return _HIP


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation holding a numpy array's bytes."""

    def __init__(self, arr=None, nbytes=None, dtype=np.float64, shape=None):
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.shape = arr.dtype, arr.shape
            nbytes = arr.nbytes
        else:
            self.dtype, self.shape = np.dtype(dtype), shape
        self.nbytes = int(nbytes)
        p = c_void_p()
        check(hip().hipMalloc(byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        if arr is not None and self.nbytes:
            self.upload(arr)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(hip().hipMemcpy(c_void_p(self.ptr), arr.ctypes.data_as(c_void_p), self.nbytes, H2D), "upload")

    def copy_from(self, other, stream=None):
        assert other.nbytes == self.nbytes
        check(hip().hipMemcpyAsync(c_void_p(self.ptr), c_void_p(other.ptr), self.nbytes, D2D,
                                   c_void_p(stream) if stream else None), "copy_from")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(hip().hipMemcpy(out.ctypes.data_as(c_void_p), c_void_p(self.ptr), self.nbytes, D2H),
                  "download")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = c_void_p()
        check(hip().hipStreamCreate(byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(c_void_p(self.handle)), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = c_void_p()
        check(hip().hipEventCreate(byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        check(hip().hipEventRecord(c_void_p(self.handle), c_void_p(stream) if stream else None), "hipEventRecord")

    def elapsed_ms(self, end):
        check(hip().hipEventSynchronize(c_void_p(end.handle)), "hipEventSynchronize")
        ms = c_float()
        check(hip().hipEventElapsedTime(byref(ms), c_void_p(self.handle), c_void_p(end.handle)),
              "hipEventElapsedTime")
        return ms.value
