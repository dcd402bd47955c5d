"""Device buffers for the GPU tests, through the HIP runtime the library
itself links (libamdhip64.so.7): float64 arrays copied in and out."""
import ctypes as C

import numpy as np

_hip = None


class DevArray:
    def __init__(self, host):
        global _hip
        if _hip is None:
            _hip = C.CDLL("libamdhip64.so.7")
        host = np.asarray(host, dtype=np.float64)
        self.shape = host.shape
        self.order = "F" if host.flags.f_contiguous and not host.flags.c_contiguous else "C"
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(8, self.nbytes))) == 0
        self.put(host)

    def put(self, host):
        host = np.asarray(host, dtype=np.float64)
        buf = np.asfortranarray(host) if self.order == "F" else np.ascontiguousarray(host)
        assert buf.nbytes == self.nbytes
        assert _hip.hipMemcpy(self.ptr, buf.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def get(self):
        out = np.empty(self.shape, order=self.order)
        assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
