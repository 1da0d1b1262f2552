"""Device memory for GPU tests that call the C ABI with device pointers: allocated, filled and
read through the same HIP runtime the native library loads (libamdhip64.so.7 from ROCm), so
that one runtime owns the device in the process."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from photonoxide_amd import _native  # noqa: F401  (loads the library and its runtime first)
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: HIP error {rc}")


class Buffer:
    """nbytes of device memory, freed with the object."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        check(hip().hipMalloc(C.byref(self.ptr), max(1, nbytes)), "hipMalloc")

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy to the device")
        return b

    def read(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy to the host")
        return out

    @property
    def address(self):
        return self.ptr.value

    def __del__(self):
        if self.ptr and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
