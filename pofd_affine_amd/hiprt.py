"""Minimal ctypes access to the HIP runtime: device buffers, streams and
events for callers that hand device pointers to the C ABI."""
import ctypes as C
import os

_rt = None


def rt():
    global _rt
    if _rt is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                            "libamdhip64.so")
        _rt = C.CDLL(path)
        _rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p,
                                            C.c_void_p]
        _rt.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _rt.hipEventSynchronize.argtypes = [C.c_void_p]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                  C.c_int]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _rt


def check(err, what):
    if err != 0:
        raise RuntimeError("%s failed with hipError %d" % (what, err))


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        check(rt().hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)),
              "hipMalloc")
        self.nbytes = nbytes

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes
        check(rt().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p),
                             arr.nbytes, 1), "hipMemcpy H2D")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        check(rt().hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptr,
                             arr.nbytes, 2), "hipMemcpy D2H")

    def __del__(self):
        if self.ptr and _rt is not None:
            _rt.hipFree(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self):
        self.ptr = C.c_void_p()
        check(rt().hipStreamCreate(C.byref(self.ptr)), "hipStreamCreate")

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")


class Event:
    def __init__(self):
        self.ptr = C.c_void_p()
        check(rt().hipEventCreate(C.byref(self.ptr)), "hipEventCreate")

    def record(self, stream):
        check(rt().hipEventRecord(self.ptr, stream.ptr), "hipEventRecord")

    def elapsed_ms(self, other):
        """Milliseconds from this event to `other` (both recorded)."""
        check(rt().hipEventSynchronize(other.ptr), "hipEventSynchronize")
        ms = C.c_float()
        check(rt().hipEventElapsedTime(C.byref(ms), self.ptr, other.ptr),
              "hipEventElapsedTime")
        return ms.value


def device_synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
