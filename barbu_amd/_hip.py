"""Minimal ctypes access to the HIP runtime for timing (events, sync, streams)."""
import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip = h
    return _hip


def _ok(r, what):
    if r != 0:
        raise RuntimeError("%s failed with hipError %d" % (what, r))


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.h, stream), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float(0)
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value


def copy_to_host(dst_array, src_device_ptr):
    """Blocking device-to-host copy into a contiguous numpy array."""
    _ok(hip().hipMemcpy(C.c_void_p(dst_array.ctypes.data), src_device_ptr, dst_array.nbytes, 2), "hipMemcpy")


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")
