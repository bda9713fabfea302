"""Drop-in for the reference's native module ``scaling_elections`` (pybind11).

Reference: ``PYBIND11_MODULE(scaling_elections, m)``, scaling_elections.cu:991-1038:
``compute_strongest_paths`` (:863-989), ``log_gpus`` (:996-1016), ``reduce`` (:1019-1032).
Argument names, keyword-only flags, defaults, result dtype and the RuntimeError texts
are the reference's; the computation runs through the C ABI of the HIP library.
"""
from __future__ import annotations

import contextlib
import ctypes
import numpy as np

from . import _lib

_SIG = ("compute_strongest_paths(): incompatible function arguments. The following argument "
        "types are supported:\n    1. (preferences: numpy.ndarray[numpy.uint32], *, "
        "allow_tma: bool = False, allow_gpu: bool = False, tile_size: int = 0) -> "
        "numpy.ndarray[numpy.uint32]")


def _as_u32_c(preferences) -> np.ndarray:
    """Input conversion of ``py::array_t<uint32_t, py::array::c_style>`` without forcecast
    (scaling_elections.cu:864): arrays convert only by safe casts (u8/u16 -> u32), other
    dtypes are rejected with TypeError; non-contiguous arrays are copied; non-array
    sequences are converted element-wise."""
    if isinstance(preferences, np.ndarray):
        if preferences.dtype == np.uint32:
            return np.ascontiguousarray(preferences)
        if np.can_cast(preferences.dtype, np.uint32, casting="safe"):
            return np.ascontiguousarray(preferences, dtype=np.uint32)
        raise TypeError(_SIG)
    try:
        return np.ascontiguousarray(np.array(preferences, dtype=np.uint32))
    except (TypeError, ValueError, OverflowError) as e:
        raise TypeError(_SIG) from e


@contextlib.contextmanager
def _sigint_scope():
    """SIGINT sets the library's stop flag while a closure runs, as the reference's module
    handler does (scaling_elections.cu:98-100, :993); the previous handler comes back after."""
    L = _lib.lib()
    installed = L.se_install_sigint() == 0
    try:
        yield
    finally:
        if installed:
            L.se_restore_sigint()


def compute_strongest_paths(preferences, *, allow_tma: bool = False, allow_gpu: bool = False,
                            tile_size: int = 0) -> np.ndarray:
    """Strongest-path (widest-path) matrix of a pairwise preference matrix.

    Reference: scaling_elections.cu:863-989.  ``allow_gpu=True`` runs the gfx950 HIP
    closure (tile_size must be 4, 8, 16 or 32 -- 0 is rejected exactly as in the
    reference; the result does not depend on it); ``allow_gpu=False`` runs the CPU branch
    with the reference's tile rules.  ``allow_tma`` is accepted and ignored (AMD GPUs
    have no TMA, scaling_elections.cu:555-558).  Returns a new (n, n) uint32 array.
    """
    prefs = _as_u32_c(preferences)
    if prefs.ndim != 2:
        raise RuntimeError("Number of dimensions must be two")
    if prefs.shape[0] != prefs.shape[1]:
        raise RuntimeError("Preferences matrix must be square")
    tile = int(tile_size)
    if tile < 0 or tile >= 1 << 64:  # std::size_t in the reference signature
        raise TypeError(_SIG)
    tile = min(tile, 0xFFFFFFFF)  # larger tiles fail the same checks as 2^32 - 1
    n = prefs.shape[0]
    result = np.empty((n, n), dtype=np.uint32)
    L = _lib.lib()
    src = _lib.ptr(prefs) if n else None
    dst = _lib.ptr(result) if n else None
    with _sigint_scope():
        if allow_gpu:
            flags = _lib.SE_FLAG_ALLOW_TMA if allow_tma else 0
            rc = L.se_closure(src, n, n, dst, tile, flags, None)
        else:
            rc = L.se_closure_cpu(src, n, n, dst, tile, 0)
    _lib.check(rc)
    return result


def compute_strongest_paths_serial(preferences) -> np.ndarray:
    """Serial CPU closure (the reference CLI's validation baseline, scaling_elections.py:91-125)."""
    prefs = _as_u32_c(preferences)
    if prefs.ndim != 2:
        raise RuntimeError("Number of dimensions must be two")
    if prefs.shape[0] != prefs.shape[1]:
        raise RuntimeError("Preferences matrix must be square")
    n = prefs.shape[0]
    result = np.empty((n, n), dtype=np.uint32)
    L = _lib.lib()
    with _sigint_scope():
        rc = L.se_closure_cpu_serial(_lib.ptr(prefs) if n else None, n, n, _lib.ptr(result) if n else None)
    _lib.check(rc)
    return result


def strongest_paths_timed(preferences, flags: int = 0) -> tuple:
    """GPU closure of a host matrix returning (paths, timing dict) -- for benchmarks."""
    prefs = _as_u32_c(preferences)
    n = prefs.shape[0]
    result = np.empty((n, n), dtype=np.uint32)
    t = _lib.se_timing()
    rc = _lib.lib().se_closure(_lib.ptr(prefs) if n else None, n, n, _lib.ptr(result) if n else None,
                               32, flags | _lib.SE_FLAG_TIME_KERNELS, ctypes.byref(t))
    _lib.check(rc)
    return result, t.as_dict()


def device_count() -> int:
    c = ctypes.c_int(0)
    rc = _lib.lib().se_device_count(ctypes.byref(c))
    if rc != 0:
        return 0
    return int(c.value)


def device_info(device: int = 0) -> dict:
    L = _lib.lib()
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    cus, mem = ctypes.c_int(0), ctypes.c_size_t(0)
    major, minor = ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(L.se_device_info(device, name, 256, ctypes.byref(cus), ctypes.byref(mem),
                                ctypes.byref(major), ctypes.byref(minor), arch, 256))
    return {"name": name.value.decode(), "compute_units": cus.value, "global_mem": mem.value,
            "major": major.value, "minor": minor.value, "arch": arch.value.decode()}


def log_gpus() -> None:
    """Prints every visible device (scaling_elections.cu:996-1016)."""
    c = ctypes.c_int(0)
    rc = _lib.lib().se_device_count(ctypes.byref(c))
    if rc != 0 or c.value <= 0:
        raise RuntimeError("No HIP devices available")
    for i in range(c.value):
        info = device_info(i)
        print(f"Device {i}: {info['name']}")
        print(f"\tCUs: {info['compute_units']}")
        print(f"\tGlobal mem: {info['global_mem'] / (1024 ** 3):.2f}GB")
        print(f"\tArch: {info['arch']} ({info['major']}.{info['minor']})")


def reduce(data) -> float:
    """Sum of a float32 array in sequential order -- the reference's non-CUDA ``reduce``
    (std::accumulate with a float accumulator, scaling_elections.cu:1029-1030)."""
    a = np.ascontiguousarray(data, dtype=np.float32).ravel()
    return float(_lib.lib().se_reduce_f32(_lib.ptr(a) if a.size else None, a.size))


def interrupt() -> None:
    """Asks a running closure to stop at its next round ("Stopped by signal")."""
    _lib.lib().se_interrupt()
