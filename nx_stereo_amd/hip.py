"""Device memory, streams and events through the HIP runtime libnxs_hip.so links.

A few ctypes calls into libamdhip64, loaded after the library so that both
share one runtime in the process (torch bundles a runtime of its own, which
does not mix with this one). For the benchmark and tests: the product is
the C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_hip = None

HOST_TO_DEVICE, DEVICE_TO_HOST, DEVICE_TO_DEVICE = 1, 2, 3


def runtime() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        _native.load()  # its libamdhip64 first, then the same one here
        lib = ctypes.CDLL("libamdhip64.so.7")
        P, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        lib.hipMalloc.argtypes = [ctypes.POINTER(P), sz]
        lib.hipFree.argtypes = [P]
        lib.hipMemcpy.argtypes = [P, P, sz, i]
        lib.hipMemcpyAsync.argtypes = [P, P, sz, i, P]
        lib.hipMemsetAsync.argtypes = [P, i, sz, P]
        lib.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(P), ctypes.c_uint]
        lib.hipStreamDestroy.argtypes = [P]
        lib.hipStreamSynchronize.argtypes = [P]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        lib.hipEventDestroy.argtypes = [P]
        lib.hipEventRecord.argtypes = [P, P]
        lib.hipEventSynchronize.argtypes = [P]
        lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        lib.hipGetErrorString.argtypes = [i]
        lib.hipGetErrorString.restype = ctypes.c_char_p
        _hip = lib
    return _hip


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise _native.NativeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


class Buffer:
    """A device allocation; `ptr` is its address."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def of(cls, array: np.ndarray) -> "Buffer":
        array = np.ascontiguousarray(array)
        buf = cls(array.nbytes)
        buf.upload(array)
        return buf

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array)
        assert array.nbytes <= self.nbytes
        check(runtime().hipMemcpy(ctypes.c_void_p(self.ptr), array.ctypes.data_as(ctypes.c_void_p),
                                  array.nbytes, HOST_TO_DEVICE), "hipMemcpy")

    def download(self, shape, dtype) -> np.ndarray:
        return copy_to_host(self.ptr, shape, dtype)

    def free(self) -> None:
        if self.ptr:
            runtime().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_to_host(ptr: int, shape, dtype) -> np.ndarray:
    out = np.empty(shape, dtype)
    check(runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes,
                              DEVICE_TO_HOST), "hipMemcpy")
    return out


class Stream:
    """A non-blocking stream at the default priority."""

    def __init__(self):
        s = ctypes.c_void_p()
        check(runtime().hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")
        self.handle = s.value

    def synchronize(self) -> None:
        check(runtime().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            runtime().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = 0


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream) -> None:
        check(runtime().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream.handle)),
              "hipEventRecord")

    def synchronize(self) -> None:
        check(runtime().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        check(runtime().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle),
                                            ctypes.c_void_p(later.handle)), "hipEventElapsedTime")
        return ms.value


def synchronize() -> None:
    check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
