"""Minimal HIP runtime access through ctypes: device buffers, copies and
events (for callers that keep rho/potential resident on the device without
going through PyTorch)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        _hip = L
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP status %d" % (what, rc))


class DeviceArray:
    """A float64 device buffer with the shape of a host array."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _chk(hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        _chk(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 1), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape)
        _chk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    _chk(hip().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value


class Timer:
    """HIP events recorded on a given stream (a hipStream_t as int/None)."""

    def __init__(self):
        self.e = []
        for _ in range(2):
            e = C.c_void_p()
            _chk(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
            self.e.append(e.value)

    def start(self, stream):
        _chk(hip().hipEventRecord(C.c_void_p(self.e[0]), C.c_void_p(stream)), "hipEventRecord")

    def stop(self, stream):
        _chk(hip().hipEventRecord(C.c_void_p(self.e[1]), C.c_void_p(stream)), "hipEventRecord")

    def ms(self):
        _chk(hip().hipEventSynchronize(C.c_void_p(self.e[1])), "hipEventSynchronize")
        v = C.c_float()
        _chk(hip().hipEventElapsedTime(C.byref(v), C.c_void_p(self.e[0]), C.c_void_p(self.e[1])), "elapsed")
        return float(v.value)
