"""Minimal HIP runtime plumbing over ctypes: device buffers, streams, events.

Only what the tests, the benchmark and the serving driver need to hand
libcraw device pointers and a hipStream_t; all path work happens in
libcraw.so's kernels.
"""
import ctypes
import os

import numpy as np

_HIP_CANDIDATES = ["libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"]
H2D, D2H, D2D = 1, 2, 3

_hip = None


def hip():
    global _hip
    if _hip is None:
        err = None
        for name in _HIP_CANDIDATES:
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError as e:
                err = e
        if _hip is None:
            raise RuntimeError("HIP runtime not found: %s" % err)
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int, ctypes.c_void_p]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                             ctypes.c_void_p]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]
    return _hip


def check(ret, what="HIP call"):
    if ret != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(ret).decode()))


def device_count():
    n = ctypes.c_int(0)
    try:
        if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
            return 0
    except RuntimeError:
        return 0
    return n.value


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """A hipMalloc'd buffer; freed on close() or garbage collection."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
              "hipMemcpy H2D")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H),
              "hipMemcpy D2H")
        return out

    def copy_from(self, other, nbytes):
        """Device-to-device copy (HIP's own copy kernel)."""
        assert nbytes <= self.nbytes and nbytes <= other.nbytes
        check(hip().hipMemcpy(self.ptr, other.ptr, nbytes, D2D), "hipMemcpy D2D")

    def zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle.value:
            hip().hipStreamDestroy(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(hip().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        check(hip().hipEventRecord(self.handle, stream.handle if stream else None),
              "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        check(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle),
              "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle.value:
                hip().hipEventDestroy(self.handle)
        except Exception:
            pass
