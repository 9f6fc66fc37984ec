"""Minimal ctypes access to the HIP runtime: device buffers and events.

Used by bench.py to keep outputs resident in HBM and to time work on the
engine's own stream (torch events would only see torch's current stream).
"""
import ctypes
import ctypes.util
import os

_hip = None


def lib():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        _hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _ok(lib().hipMalloc(ctypes.byref(self.ptr), nbytes), "hipMalloc")
        self.nbytes = nbytes

    def from_host(self, arr, offset=0):
        """Copy numpy array arr to byte offset of the buffer."""
        _ok(lib().hipMemcpy(self.ptr.value + offset, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def to_host(self, arr, nbytes=None, offset=0):
        """Copy nbytes (default: arr.nbytes) starting at byte offset into numpy array arr."""
        n = arr.nbytes if nbytes is None else nbytes
        _ok(lib().hipMemcpy(arr.ctypes.data, self.ptr.value + offset, n, 2), "hipMemcpy D2H")

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _ok(lib().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream):
        _ok(lib().hipEventRecord(self.h, stream), "hipEventRecord")

    def synchronize(self):
        _ok(lib().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = ctypes.c_float()
        _ok(lib().hipEventElapsedTime(ctypes.byref(t), self.h, end.h), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        if self.h:
            lib().hipEventDestroy(self.h)


def device_attribute(attr, device=0):
    v = ctypes.c_int()
    _ok(lib().hipDeviceGetAttribute(ctypes.byref(v), attr, device), "hipDeviceGetAttribute")
    return v.value


# hipDeviceAttribute_t values (hip_runtime_api.h)
ATTR_CLOCK_RATE_KHZ = 5  # hipDeviceAttributeClockRate
ATTR_CU_COUNT = 63       # hipDeviceAttributeMultiprocessorCount
