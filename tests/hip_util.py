"""Device buffers for tests that call entry points on caller-owned device
arrays, taken from the HIP runtime the library itself is linked against (a
second HIP runtime in the process cannot open the device)."""
import ctypes

import numpy as np


class DeviceBuffers:
    """The caller's device buffers, from the HIP runtime the library itself is
    linked against (a second runtime in the process cannot open the device)."""

    def __init__(self):
        import twochnrg_amd as t
        t.lib()
        self.rt = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipSetDevice.argtypes = [ctypes.c_int]
        assert self.rt.hipSetDevice(0) == 0

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(8, a.nbytes)) == 0
        if a.nbytes:
            assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, n):
        a = np.empty(n)
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0      # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self.rt.hipFree(p)
