"""Device buffers and streams through the HIP runtime the engine links
(libamdhip64 from /opt/rocm), for tests that hand device pointers across
the C ABI.  Only plumbing: allocation, copies, streams."""
import ctypes
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                            "libamdhip64.so")
        _hip = ctypes.CDLL(path)
        for f in ("hipMalloc",):
            getattr(_hip, f).argtypes = [ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_int]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: hipError %d" % (what, rc))


class DeviceBuffer:
    """nbytes of device memory; to_device / to_host copy numpy arrays."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.to_device(a)
        return b

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1),
            "hipMemcpy H2D")

    def to_host(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2),
            "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.handle)),
            "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()
