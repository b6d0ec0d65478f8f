"""A few HIP runtime calls through ctypes: device buffers, copies and events, for timing
the stepper on its own stream without going through torch."""

from __future__ import annotations

import ctypes

import numpy as np

_hip = None


def runtime():
    """The HIP runtime the wave-solver library is bound to: loaded by its soname once the
    library is, so that the process holds one runtime (its streams, events and buffers are
    then the library's own)."""
    global _hip
    if _hip is None:
        from acoustic_spaces_amd import native

        native.load()
        lib = ctypes.CDLL("libamdhip64.so.7")
        P = ctypes.c_void_p
        lib.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
        lib.hipFree.argtypes = [P]
        lib.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipStreamSynchronize.argtypes = [P]
        lib.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        lib.hipEventDestroy.argtypes = [P]
        lib.hipEventRecord.argtypes = [P, P]
        lib.hipEventSynchronize.argtypes = [P]
        lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        lib.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        lib.hipGetErrorString.restype = ctypes.c_char_p
        lib.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = lib
    return _hip


def _check(code):
    if code != 0:
        raise RuntimeError(runtime().hipGetErrorString(code).decode())


def device_count() -> int:
    n = ctypes.c_int(0)
    if runtime().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceArray:
    """A device buffer holding a copy of a host array."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = max(host.nbytes, 4)
        self.dtype = host.dtype
        self.shape = host.shape
        p = ctypes.c_void_p()
        _check(runtime().hipMalloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        if host.nbytes:
            _check(runtime().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1))  # host to device

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2))  # device to host
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(runtime().hipEventCreate(ctypes.byref(e)))
        self.e = e.value

    def record(self, stream: int):
        _check(runtime().hipEventRecord(self.e, stream))

    def elapsed_ms(self, later: "Event") -> float:
        _check(runtime().hipEventSynchronize(later.e))
        ms = ctypes.c_float()
        _check(runtime().hipEventElapsedTime(ctypes.byref(ms), self.e, later.e))
        return float(ms.value)

    def __del__(self):
        try:
            runtime().hipEventDestroy(self.e)
        except Exception:
            pass


def synchronize():
    _check(runtime().hipDeviceSynchronize())
