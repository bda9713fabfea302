"""ctypes bindings over the C ABI (include/scaling_elections_amd.h).

The library is built in-tree (``scalingelections_amd/libscaling_elections_amd.so``) by
``scalingelections_amd.build``.  There is no fallback: when the library is missing or
does not load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscaling_elections_amd.so")

SE_OK = 0
SE_ERR_INVALID = -1
SE_ERR_TILE = -2
SE_ERR_ALLOC = -3
SE_ERR_DEVICE = -4
SE_ERR_COPY = -5
SE_ERR_NO_DEVICE = -6
SE_ERR_INTERRUPTED = -7
SE_ERR_RANGE = -8

SE_FLAG_AUTO_WIDTH = 0
SE_FLAG_FORCE_U32 = 1
SE_FLAG_FORCE_U16 = 2
SE_FLAG_TIME_KERNELS = 4
SE_FLAG_ALLOW_TMA = 8
SE_FLAG_PIVOT32 = 16
SE_FLAG_PIVOT64 = 32
SE_FLAG_NO_F16 = 128
SE_FLAG_NO_LOOKAHEAD = 256
SE_FLAG_LOOKAHEAD = 512
SE_FLAG_LOOKAHEAD_FUSED = 1024

# Every symbol include/scaling_elections_amd.h declares.
EXPORTED = (
    "se_last_error", "se_version", "se_device_count", "se_device_info",
    "se_closure", "se_closure_device", "se_closure_cpu", "se_closure_cpu_serial",
    "se_count", "se_count_device", "se_count_rankings",
    "se_rank", "se_rank_device", "se_reduce_f32",
    "se_interrupt", "se_clear_interrupt", "se_install_sigint", "se_restore_sigint",
    "se_batch_device", "se_synth_mallows", "se_stream_release",
    "se_tune_closure", "se_tune_closure_get", "se_tune_batch", "se_tune_batch_get",
    "se_positions16_device",
)

SE_TUNE_F16 = 0
SE_TUNE_U16 = 1
SE_TUNE_U32 = 2
SE_TUNE_MAX_ENTRIES = 64

SE_BATCH_PER_LAUNCH = 64
SE_BATCH_ONE_GROUP = 1024
SE_BATCH_POS16 = 2048


class se_batch_timing(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("count_ms", ctypes.c_double),
        ("closure_ms", ctypes.c_double),
        ("elections", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("cells", ctypes.c_uint64),
        ("pair_votes", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class se_tune_entry(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("max_side", ctypes.c_uint32),
        ("pivot", ctypes.c_uint32),
        ("lookahead", ctypes.c_uint32),
    ]


class se_timing(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("init_ms", ctypes.c_double),
        ("diag_ms", ctypes.c_double),
        ("panel_ms", ctypes.c_double),
        ("indep_ms", ctypes.c_double),
        ("output_ms", ctypes.c_double),
        ("rounds", ctypes.c_uint32),
        ("pivot", ctypes.c_uint32),
        ("width_bits", ctypes.c_uint32),
        ("padded_n", ctypes.c_uint32),
        ("indep_cells", ctypes.c_uint64),
        ("f16_ordered", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class SEError(RuntimeError):
    """RuntimeError carrying the library status code."""

    def __init__(self, msg: str, code: int):
        super().__init__(msg)
        self.code = code


_lib = None
_lock = threading.Lock()


def _declare(L):
    c = ctypes
    vp, u32, i32, sz, i64 = c.c_void_p, c.c_uint32, c.c_int32, c.c_size_t, c.c_int64
    L.se_last_error.argtypes = []
    L.se_last_error.restype = c.c_char_p
    L.se_version.argtypes = []
    L.se_version.restype = c.c_int
    L.se_device_count.argtypes = [c.POINTER(c.c_int)]
    L.se_device_count.restype = c.c_int
    L.se_device_info.argtypes = [c.c_int, c.c_char_p, sz, c.POINTER(c.c_int), c.POINTER(sz),
                                 c.POINTER(c.c_int), c.POINTER(c.c_int), c.c_char_p, sz]
    L.se_device_info.restype = c.c_int
    L.se_closure.argtypes = [vp, u32, sz, vp, u32, u32, c.POINTER(se_timing)]
    L.se_closure.restype = c.c_int
    L.se_closure_device.argtypes = [vp, u32, sz, vp, sz, u32, vp, c.POINTER(se_timing)]
    L.se_closure_device.restype = c.c_int
    L.se_closure_cpu.argtypes = [vp, u32, sz, vp, u32, c.c_int]
    L.se_closure_cpu.restype = c.c_int
    L.se_closure_cpu_serial.argtypes = [vp, u32, sz, vp]
    L.se_closure_cpu_serial.restype = c.c_int
    L.se_count.argtypes = [vp, u32, u32, vp, sz]
    L.se_count.restype = c.c_int
    L.se_count_device.argtypes = [vp, u32, u32, vp, sz, vp, c.POINTER(c.c_double)]
    L.se_count_device.restype = c.c_int
    L.se_count_rankings.argtypes = [vp, vp, u32, u32, vp, sz]
    L.se_count_rankings.restype = c.c_int
    L.se_rank.argtypes = [vp, u32, sz, vp, vp]
    L.se_rank.restype = c.c_int
    L.se_rank_device.argtypes = [vp, u32, sz, vp, vp, vp]
    L.se_rank_device.restype = c.c_int
    L.se_reduce_f32.argtypes = [vp, sz]
    L.se_reduce_f32.restype = c.c_float
    L.se_interrupt.argtypes = []
    L.se_interrupt.restype = None
    L.se_clear_interrupt.argtypes = []
    L.se_clear_interrupt.restype = None
    L.se_install_sigint.argtypes = []
    L.se_install_sigint.restype = c.c_int
    L.se_restore_sigint.argtypes = []
    L.se_restore_sigint.restype = None
    L.se_batch_device.argtypes = [u32, vp, vp, vp, vp, vp, u32, vp, c.POINTER(se_batch_timing)]
    L.se_batch_device.restype = c.c_int
    L.se_synth_mallows.argtypes = [u32, vp, vp, vp, vp, c.c_uint64, vp, c.c_int]
    L.se_synth_mallows.restype = c.c_int
    L.se_stream_release.argtypes = [vp]
    L.se_stream_release.restype = c.c_int
    L.se_positions16_device.argtypes = [vp, sz, vp, vp]
    L.se_positions16_device.restype = c.c_int
    L.se_tune_closure.argtypes = [c.POINTER(se_tune_entry), u32]
    L.se_tune_closure.restype = c.c_int
    L.se_tune_closure_get.argtypes = [c.POINTER(se_tune_entry), u32, c.POINTER(u32)]
    L.se_tune_closure_get.restype = c.c_int
    L.se_tune_batch.argtypes = [u32]
    L.se_tune_batch.restype = c.c_int
    L.se_tune_batch_get.argtypes = [c.POINTER(u32)]
    L.se_tune_batch_get.restype = c.c_int
    del i32, i64


def lib():
    """The loaded library; raises RuntimeError when it is not built or fails to load."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"HIP library not built: {LIB_PATH} is missing "
                        "(run `python -m scalingelections_amd.build`)")
                L = ctypes.CDLL(LIB_PATH)
                _declare(L)
                _lib = L
    return _lib


def last_error() -> str:
    msg = lib().se_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(rc: int) -> None:
    if rc != SE_OK:
        raise SEError(last_error() or f"scaling_elections_amd error {rc}", rc)


def ptr(a) -> int:
    """Address of a numpy array's data (or 0 for None)."""
    if a is None:
        return 0
    return a.ctypes.data
