"""Device buffers for tests of the device-pointer entry points, allocated
through the HIP runtime libbdstep.so itself loads (ROCm's libamdhip64).
torch ships its own HIP runtime; a pointer from one runtime is not valid in
the other, so these tests keep torch out of the process's device work."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from brownpak_amd import load_library
        load_library()                      # the library's runtime, already mapped
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevArray:
    """An FP64 device array with host copies in and out."""

    def __init__(self, host):
        h = np.ascontiguousarray(host, dtype=np.float64)
        self.shape = h.shape
        self.nbytes = h.nbytes
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
