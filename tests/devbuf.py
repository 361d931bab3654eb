"""Device memory for the GPU tests, from the HIP runtime the library links."""
import numpy as np


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself links (a second
    runtime, such as torch's, may not share the device with it in one process)."""
    _hip = None

    def __init__(self, nbytes):
        import ctypes
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        assert DeviceBuffer._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        import ctypes
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert DeviceBuffer._hip.hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, dtype, count):
        import ctypes
        a = np.empty(count, dtype)
        assert DeviceBuffer._hip.hipMemcpy(ctypes.c_void_p(a.ctypes.data), self.ptr, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        DeviceBuffer._hip.hipFree(self.ptr)
