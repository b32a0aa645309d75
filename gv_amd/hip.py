"""Minimal ctypes access to the HIP runtime: device buffers and events.

Plumbing for callers that hand device pointers to the C ABI (benchmarks,
tests) without pulling in a framework.  Events are recorded on a given
stream (e.g. a graph handle's own stream from GraphVine.stream()).
"""
import ctypes

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        P = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
        L.hipFree.argtypes = [P]
        L.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
        L.hipMemcpyAsync.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int, P]
        L.hipMemset.argtypes = [P, ctypes.c_int, ctypes.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        L.hipEventRecord.argtypes = [P, P]
        L.hipEventSynchronize.argtypes = [P]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        L.hipEventDestroy.argtypes = [P]
        L.hipStreamSynchronize.argtypes = [P]
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        L.hipGetErrorString.restype = ctypes.c_char_p
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A device buffer with a numpy dtype and length."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(len(a), a.dtype)
        if d.nbytes:
            _ok(hip().hipMemcpy(d.ptr, a.ctypes.data_as(ctypes.c_void_p), d.nbytes, H2D), "H2D")
        return d

    def view(self, start, n):
        """Non-owning view of entries [start, start+n) (valid while self lives)."""
        start, n = int(start), int(n)
        if start < 0 or n < 0 or start + n > self.n:
            raise ValueError("view out of range")
        v = DeviceArray.__new__(DeviceArray)
        v.dtype = self.dtype
        v.n = n
        v.ptr = ctypes.c_void_p((self.ptr.value or 0) + start * self.dtype.itemsize)
        v._base = self
        return v

    def upload(self, a):
        """Overwrite the first len(a) entries from a host array (synchronous;
        the caller orders it after any stream work that reads the buffer)."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if len(a) > self.n:
            raise ValueError("host array larger than the device buffer")
        if len(a):
            _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                len(a) * self.dtype.itemsize, H2D), "H2D")

    def copy_from_async(self, src, stream):
        """Enqueue a device-to-device copy of src (a DeviceArray or view of
        the same dtype, at most len(self) entries) into the first entries on
        the given stream."""
        if src.dtype != self.dtype or src.n > self.n:
            raise ValueError("source must have this dtype and fit")
        if src.n:
            _ok(hip().hipMemcpyAsync(self.ptr, src.ptr, src.nbytes, D2D,
                                     ctypes.c_void_p(stream)), "D2D async")

    def numpy(self, n=None):
        n = self.n if n is None else int(n)
        out = np.empty(n, self.dtype)
        if n:
            _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                n * self.dtype.itemsize, D2H), "D2H")
        return out

    def free(self):
        if getattr(self, "_base", None) is not None:
            self._base = None  # a view owns nothing
            self.ptr = ctypes.c_void_p()
            return
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream):
        _ok(hip().hipEventRecord(self.e, ctypes.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(t), self.e, end.e), "hipEventElapsedTime")
        return float(t.value)


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
