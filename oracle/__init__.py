"""CPU oracle for the Schulze path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this package.  It is the checker the product is compared against, never the
thing measured or shipped: ``scalingelections_amd`` does not import it.

Bindings (ctypes) over ``oracle/lib/liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``) and, where it was built, ``oracle/_ref/libref.so`` (the
reference's own OpenMP CPU closure compiled from its source).  Each wrapper names
the reference lines its C function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

_lib = None
_ref = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(reference: Optional[str] = None) -> None:
    """Compile the oracle (and the reference CPU closure when its source is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    ref = reference or "/root/reference"
    if os.path.exists(os.path.join(ref, "scaling_elections.cu")):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REF={ref}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64 = ctypes.c_int64
        L.oracle_p0.argtypes = [_u32p, i64, i64, _u32p, i64]
        L.oracle_p0.restype = None
        L.oracle_closure_serial.argtypes = [_u32p, i64, i64, _u32p]
        L.oracle_closure_serial.restype = None
        L.oracle_closure_tiled.argtypes = [_u32p, i64, i64, _u32p, i64, ctypes.c_int, i64]
        L.oracle_closure_tiled.restype = ctypes.c_int
        L.oracle_widest_row.argtypes = [_u32p, i64, i64, i64, _u32p]
        L.oracle_widest_row.restype = ctypes.c_int
        L.oracle_populate.argtypes = [_u32p, i64, _i64p, i64]
        L.oracle_populate.restype = None
        L.oracle_build_prefs.argtypes = [_i64p, _i64p, i64, i64, _u32p]
        L.oracle_build_prefs.restype = ctypes.c_int
        L.oracle_wins.argtypes = [_u32p, i64, _i64p]
        L.oracle_wins.restype = None
        L.oracle_ranking.argtypes = [_i64p, i64, _i64p]
        L.oracle_ranking.restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(_REF_PATH)
        R.ref_closure_openmp.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32]
        R.ref_closure_openmp.restype = ctypes.c_int
        _ref = R
    return _ref


def _square_u32(d) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.uint32)
    assert d.ndim == 2 and d.shape[0] == d.shape[1]
    return d


def p0(d) -> np.ndarray:
    """Strongest-path init, scaling_elections.py:103-112."""
    d = _square_u32(d)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    lib().oracle_p0(d, n, n, out, n)
    return out


def closure_serial(d) -> np.ndarray:
    """compute_strongest_paths_numba_serial, scaling_elections.py:91-125."""
    d = _square_u32(d)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    lib().oracle_closure_serial(d, n, n, out)
    return out


def closure_tiled(d, tile: int = 16, threads: int = 0, max_rounds: int = -1) -> np.ndarray:
    """compute_strongest_paths_numba_parallel, scaling_elections.py:162-271 (zero padded)."""
    d = _square_u32(d)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    rc = lib().oracle_closure_tiled(d, n, n, out, tile, threads, max_rounds)
    if rc != 0:
        raise MemoryError("oracle_closure_tiled failed")
    return out


def widest_row(p0m, src: int) -> np.ndarray:
    """Per-source widest paths by modified Dijkstra (independent of Floyd-Warshall)."""
    p0m = _square_u32(p0m)
    n = p0m.shape[0]
    out = np.empty(n, dtype=np.uint32)
    if lib().oracle_widest_row(p0m, n, n, int(src), out) != 0:
        raise MemoryError("oracle_widest_row failed")
    return out


def populate(prefs: np.ndarray, ranking) -> None:
    """populate_preferences_from_ranking, scaling_elections.py:37-48 (in place)."""
    assert prefs.dtype == np.uint32 and prefs.flags.c_contiguous
    r = np.ascontiguousarray(ranking, dtype=np.int64)
    lib().oracle_populate(prefs, prefs.shape[0], r, r.shape[0])


def build_prefs(rankings: Sequence) -> np.ndarray:
    """build_pairwise_preferences, scaling_elections.py:51-88."""
    rs = [np.ascontiguousarray(r, dtype=np.int64).ravel() for r in rankings]
    n = 1
    for r in rs:
        n = max(n, int(np.max(r)) + 1)
    offsets = np.zeros(len(rs) + 1, dtype=np.int64)
    for i, r in enumerate(rs):
        offsets[i + 1] = offsets[i] + r.shape[0]
    flat = np.concatenate(rs) if rs else np.zeros(1, dtype=np.int64)
    flat = np.ascontiguousarray(flat, dtype=np.int64)
    out = np.empty((n, n), dtype=np.uint32)
    if lib().oracle_build_prefs(flat, offsets, len(rs), n, out) != 0:
        raise MemoryError("oracle_build_prefs failed")
    return out


def wins(p) -> np.ndarray:
    """Win counts, scaling_elections.py:287-290."""
    p = _square_u32(p)
    n = p.shape[0]
    out = np.empty(n, dtype=np.int64)
    lib().oracle_wins(p, n, out)
    return out


def ranking(p) -> np.ndarray:
    """Stable descending ranking by wins, scaling_elections.py:292."""
    w = wins(p)
    n = w.shape[0]
    out = np.empty(n, dtype=np.int64)
    lib().oracle_ranking(w, n, out)
    return out


def winner_and_ranking(candidates: list, p):
    """get_winner_and_ranking, scaling_elections.py:274-296."""
    order = ranking(p)
    ranked = [candidates[int(i)] for i in order]
    return ranked[0], ranked


def ref_closure_openmp(d, tile: int = 16) -> np.ndarray:
    """The reference's own compute_strongest_paths_openmp<tile, false>
    (scaling_elections.cu:738-849), compiled from its source.  n % tile must be 0."""
    d = _square_u32(d)
    n = d.shape[0]
    out = np.zeros((n, n), dtype=np.uint32)
    if ref_lib().ref_closure_openmp(d, n, n, out, tile) != 0:
        raise ValueError("reference OpenMP closure needs n % tile == 0 and tile in 4..64")
    return out
