"""Minimal device-memory and stream plumbing over the HIP runtime the library links.

Callers of the *_device entry points own their device buffers and streams; this module
provides them without a second runtime in the process (PyTorch wheels bundle their own
libamdhip64, which cannot share the device with the system runtime the library uses).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def hip():
    """The HIP runtime already loaded by the library (resolved by soname)."""
    global _hip
    if _hip is None:
        _lib.lib()  # loads libamdhip64 as a dependency first
        H = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        H.hipMemsetAsync.argtypes = [vp, ctypes.c_int, sz, vp]
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipEventDestroy.argtypes = [vp]
        H.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemcpyAsync", "hipMemsetAsync", "hipStreamCreate",
                  "hipStreamCreateWithFlags",
                  "hipStreamDestroy", "hipStreamSynchronize", "hipDeviceSynchronize", "hipEventCreate",
                  "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime", "hipEventDestroy",
                  "hipMemGetInfo"):
            getattr(H, f).restype = ctypes.c_int
        _hip = H
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceBuffer:
    """A hipMalloc'd buffer (freed on close / garbage collection)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")

    def upload_async(self, a: np.ndarray, stream: "Stream", offset: int = 0) -> None:
        """Queues a host-to-device copy on `stream`; `a` must stay alive until the stream
        has passed the copy."""
        assert a.flags.c_contiguous and offset + a.nbytes <= self.nbytes
        _ok(hip().hipMemcpyAsync(self.ptr + offset, a.ctypes.data, a.nbytes, hipMemcpyHostToDevice, stream.handle),
            "hipMemcpyAsync H2D")

    def download_async(self, out: np.ndarray, stream: "Stream", offset: int = 0) -> None:
        assert out.flags.c_contiguous and offset + out.nbytes <= self.nbytes
        _ok(hip().hipMemcpyAsync(out.ctypes.data, self.ptr + offset, out.nbytes, hipMemcpyDeviceToHost, stream.handle),
            "hipMemcpyAsync D2H")

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    def close(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


class Stream:
    """A HIP stream at the default priority.  nonblocking=True: no implicit ordering with
    the null stream, so work on different streams overlaps."""

    def __init__(self, nonblocking: bool = False):
        s = ctypes.c_void_p()
        if nonblocking:
            _ok(hip().hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")
        else:
            _ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self) -> None:
        if self.handle:
            _lib.lib().se_stream_release(self.handle)  # scratch the library keeps for this stream
            hip().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: "Stream" = None) -> None:
        _ok(hip().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                hip().hipEventDestroy(self.handle)
        except Exception:  # noqa: BLE001
            pass


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _ok(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return int(free.value), int(total.value)


def device_synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
