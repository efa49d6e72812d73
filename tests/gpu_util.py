"""Device-memory helpers for GPU tests: allocate/copy through libdsmc's own
C ABI (dsmc_device_alloc, dsmc_memcpy_*), so the tests exercise the shipped
library and need no other device allocator."""
import ctypes as C

import numpy as np


class DeviceArray:
    def __init__(self, lib, host: np.ndarray):
        self.lib = lib
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = C.c_void_p()
        st = lib.dsmc_device_alloc(C.byref(self.ptr), max(self.nbytes, 16))
        assert st == 0, lib.dsmc_last_error(None)
        if self.nbytes:
            st = lib.dsmc_memcpy_htod(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.nbytes)
            assert st == 0, lib.dsmc_last_error(None)

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        if self.nbytes:
            st = self.lib.dsmc_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
            assert st == 0, self.lib.dsmc_last_error(None)
        return out

    def free(self):
        if self.ptr:
            self.lib.dsmc_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()
