"""Device buffers for GPU tests that pass device pointers through the C-ABI
(arrays_on_device = 1)."""
import ctypes

import numpy as np


class HipBuffers:
    """Device buffers from the HIP runtime libeafp.so itself links
    (/opt/rocm's libamdhip64): torch's wheel carries another HIP runtime,
    and in one process with libeafp's it has failed to find the GPU."""

    def __init__(self):
        self.ct = ctypes
        self.h = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        ptr = self.ct.c_void_p()
        assert self.h.hipMalloc(self.ct.byref(ptr), self.ct.c_size_t(nbytes)) == 0
        self.bufs.append(ptr)
        return ptr.value

    def put(self, a):
        a = np.ascontiguousarray(a.ravel(order="F"))
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(self.ct.c_void_p(p), self.ct.c_void_p(a.ctypes.data), self.ct.c_size_t(a.nbytes), 1) == 0
        return p

    def zeros(self, nbytes):
        p = self.alloc(nbytes)
        assert self.h.hipMemset(self.ct.c_void_p(p), 0, self.ct.c_size_t(nbytes)) == 0
        assert self.h.hipDeviceSynchronize() == 0    # libeafp's stream does not wait for the null stream
        return p

    def get(self, p, pad, dtype):
        a = np.empty(int(np.prod(pad)), dtype=dtype)
        assert self.h.hipMemcpy(self.ct.c_void_p(a.ctypes.data), self.ct.c_void_p(p), self.ct.c_size_t(a.nbytes), 2) == 0
        return a.reshape(pad, order="F")

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []
