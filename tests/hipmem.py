"""Device buffers for GPU tests through the HIP runtime that libwfamd.so itself links
(libamdhip64.so.7), the way the reference's tests use cudaMalloc/cudaMemcpy directly
(tests/test_gpu_queues.cu). torch is not used in GPU tests: it bundles a second HIP runtime that
cannot initialise in the same process.
"""
import ctypes as C

import numpy as np

_L = None
H2D, D2H = 1, 2


def hip():
    global _L
    if _L is None:
        import wfamd
        wfamd.lib()   # loads libamdhip64.so.7 first, so the name below resolves to the same copy
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipDeviceSynchronize):
            f.restype = C.c_int
        _L = L
    return _L


class DevArray:
    """A device copy of a numpy array (uint32 / int32 / float32)."""

    def __init__(self, host=None, n=None, dtype=np.uint32):
        if host is not None:
            host = np.ascontiguousarray(host)
            self.dtype, self.n = host.dtype, host.size
        else:
            self.dtype, self.n = np.dtype(dtype), n
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), max(1, self.n) * self.dtype.itemsize) == 0
        self.ptr = p.value
        if host is not None and self.n:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.n * self.dtype.itemsize, H2D) == 0

    def get(self):
        out = np.zeros(self.n, self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        if self.n:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * self.dtype.itemsize, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        if _L is not None:
            self.free()


def sync():
    assert hip().hipDeviceSynchronize() == 0
