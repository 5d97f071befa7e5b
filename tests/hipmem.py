"""Device buffers for C-ABI tests through the HIP runtime the library is
linked with (libamdhip64.so.7, the ROCm of /opt/rocm).  torch carries its own
HIP runtime; buffers of one runtime are not handed to the other."""
import ctypes as C

import numpy as np

_H = None


def _hip():
    global _H
    if _H is None:
        from conftest import load_package
        load_package()._lib.lib()             # loads libmpm3d_hip.so and with it its runtime
        H = C.CDLL("libamdhip64.so.7")        # the already loaded runtime
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _H = H
    return _H


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: hipError %d" % (what, rc))


class DeviceArray:
    """a device copy of a numpy array (or an uninitialised one of a shape and dtype)"""

    def __init__(self, host=None, shape=None, dtype=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(_hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if host is not None:
            _ok(_hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def get(self):
        _ok(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        out = np.empty(self.shape, self.dtype)
        _ok(_hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _hip().hipDeviceSynchronize()
            _hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
