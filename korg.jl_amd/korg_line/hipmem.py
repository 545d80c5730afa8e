"""Device memory through the HIP runtime that libkorgline.so itself links.

PyTorch wheels bundle their own HIP runtime; two HIP runtimes in one process do not share
devices, so the device-resident form of the library manages its buffers with the system runtime
(``libamdhip64.so.7``) directly: allocation, host<->device copies, synchronisation.
"""
import ctypes

import numpy as np

from . import _lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _lib.load()  # brings libamdhip64.so.7 into the process
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(rc).decode()}")


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def read(ptr, n, dtype):
    """``n`` elements of ``dtype`` at the device address ``ptr``, copied into a new numpy array."""
    out = np.empty(int(n), dtype=dtype)
    if out.nbytes:
        _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                               out.nbytes, D2H), "hipMemcpy D2H")
    return out


class DeviceArray:
    """A device buffer holding an array of ``shape`` and ``dtype`` (C order)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) \
            else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._p = p

    @property
    def ptr(self):
        return self._p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _check(hip().hipMemcpy(d._p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
                   "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        return cls.from_numpy(np.zeros(shape, dtype=dtype))

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(hip().hipMemcpy(self._p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
               "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self._p, self.nbytes, D2H),
                   "hipMemcpy D2H")
        return out

    def free(self):
        if self._p is not None and self._p.value:
            hip().hipFree(self._p)
        self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
