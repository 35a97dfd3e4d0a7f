"""Device buffers for tests of entry points that take device pointers: the
HIP runtime's hipMalloc / hipMemcpy / hipFree through ctypes (the test is
the caller; these are its buffers)."""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one)."""

    def __init__(self, a=None, shape=None, dtype=None):
        self.dtype = np.dtype(a.dtype if a is not None else dtype)
        self.shape = tuple(a.shape if a is not None else shape)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.p), max(self.nbytes, 4)), "hipMalloc")
        if a is not None and self.nbytes:
            a = np.ascontiguousarray(a)
            _ok(hip().hipMemcpy(self.p, a.ctypes.data, self.nbytes, _H2D), "hipMemcpy")

    @property
    def ptr(self):
        return self.p.value

    def get(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data, self.p, self.nbytes, _D2H), "hipMemcpy")
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
