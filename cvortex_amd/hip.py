"""Minimal ctypes binding of the HIP runtime: device memory, copies, streams
and events.  Used by the device-resident API wrappers, the benchmark and
the tests; the library itself is C++ and does not use this module."""
import ctypes
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                             ctypes.c_void_p]
    return _hip


def check(err, what="hip"):
    if err != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(err).decode()}")


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A device allocation holding a numpy-shaped array."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 16))),
              "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(hip().hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes),
                              H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(hip().hipDeviceSynchronize(), "sync")
        check(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr,
                              ctypes.c_size_t(self.nbytes), D2H), "hipMemcpy D2H")
        return out

    def fill_zero(self):
        check(hip().hipMemset(self.ptr, 0, ctypes.c_size_t(self.nbytes)), "hipMemset")

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check(hip().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        check(hip().hipEventRecord(self.h, ctypes.c_void_p(stream) if stream else None),
              "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(hip().hipEventElapsedTime(ctypes.byref(ms), self.h, later.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.h)
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        check(hip().hipStreamCreateWithFlags(ctypes.byref(self.h), ctypes.c_uint(1)),
              "hipStreamCreateWithFlags")

    @property
    def addr(self):
        return self.h.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def __del__(self):
        try:
            hip().hipStreamSynchronize(self.h)
            hip().hipStreamDestroy(self.h)
        except Exception:
            pass


def stream_synchronize(stream):
    check(hip().hipStreamSynchronize(ctypes.c_void_p(stream)), "hipStreamSynchronize")
