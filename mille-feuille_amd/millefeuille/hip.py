"""Minimal device-buffer plumbing over the HIP runtime (ctypes).

Uses the same libamdhip64.so.7 that libmillefeuille_hip.so links, so
pointers are valid for the library's kernels.  Only what tests and the
benchmark need: allocation, copies, memset, synchronisation, events and the
device count.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib  # noqa: F401  (loads libmillefeuille_hip.so and its HIP runtime)

_hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
for name, res, args in [
    ("hipMalloc", C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    ("hipFree", C.c_int, [C.c_void_p]),
    ("hipMemcpy", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    ("hipMemset", C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    ("hipDeviceSynchronize", C.c_int, []),
    ("hipGetDeviceCount", C.c_int, [C.POINTER(C.c_int)]),
    ("hipSetDevice", C.c_int, [C.c_int]),
    ("hipGetErrorString", C.c_char_p, [C.c_int]),
    ("hipEventCreate", C.c_int, [C.POINTER(C.c_void_p)]),
    ("hipEventRecord", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hipEventSynchronize", C.c_int, [C.c_void_p]),
    ("hipEventElapsedTime", C.c_int, [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
    ("hipEventDestroy", C.c_int, [C.c_void_p]),
]:
    f = getattr(_hip, name)
    f.restype = res
    f.argtypes = args

H2D, D2H, D2D = 1, 2, 3


def _ok(fn, e):
    if e != 0:
        raise RuntimeError(f"{fn}: {_hip.hipGetErrorString(e).decode()}")


def device_count():
    n = C.c_int(0)
    e = _hip.hipGetDeviceCount(C.byref(n))
    return n.value if e == 0 else 0


def synchronize():
    _ok("hipDeviceSynchronize", _hip.hipDeviceSynchronize())


class DeviceArray:
    """A typed device allocation (freed with the object)."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(1, self.n * self.dtype.itemsize)
        p = C.c_void_p()
        _ok("hipMalloc", _hip.hipMalloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None, length=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        n = len(a) if length is None else int(length)
        d = cls(n, a.dtype)
        d.zero()
        if len(a):
            _ok("hipMemcpy", _hip.hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D))
        return d

    def zero(self):
        _ok("hipMemset", _hip.hipMemset(self.ptr, 0, self.nbytes))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _ok("hipMemcpy", _hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D))

    def get(self, n=None):
        n = self.n if n is None else n
        out = np.empty(n, dtype=self.dtype)
        if n:
            _ok("hipMemcpy", _hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H))
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        # No HIP calls while the interpreter shuts down: the runtime's own
        # exit handlers (and a profiler's) may already have run, and the
        # process teardown releases device memory anyway.
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """HIP event pair on the null stream."""

    def __init__(self):
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        _ok("hipEventCreate", _hip.hipEventCreate(C.byref(self.e0)))
        _ok("hipEventCreate", _hip.hipEventCreate(C.byref(self.e1)))

    def start(self):
        _ok("hipEventRecord", _hip.hipEventRecord(self.e0, None))

    def stop(self):
        _ok("hipEventRecord", _hip.hipEventRecord(self.e1, None))
        _ok("hipEventSynchronize", _hip.hipEventSynchronize(self.e1))
        ms = C.c_float()
        _ok("hipEventElapsedTime", _hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1))
        return ms.value

    def free(self):
        if self.e0:
            _hip.hipEventDestroy(self.e0)
            _hip.hipEventDestroy(self.e1)
            self.e0 = self.e1 = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass
