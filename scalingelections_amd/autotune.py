"""Launch-shape autotuner: measures the fastest closure and batch shapes on the current GPU.

No reference counterpart: the reference's GPU path runs whatever tile size the caller
passes (scaling_elections.cu:895-901).  Here the library picks its launch shapes from a
process-wide table (include/scaling_elections_amd.h, se_tune_closure / se_tune_batch):

* per value width (f16-ordered 16-bit, integer 16-bit, 32-bit) and zero-padded side N,
  the closure's pivot depth (32 or 64) and look-ahead mode (0: off; 1: the next pivot's
  panel overlaps the current round, its strips in a launch of their own; 2: the same in
  one launch per round);
* the batch's threshold above which elections share per-round launches instead of being
  closed by one workgroup each.

The built-in table holds what this tuner measured on MI355X.  ``tune_closure`` and
``tune_batch`` re-measure on the current device (every candidate's result is checked equal
to the others' first: no choice may change a result), ``install`` / ``save`` / ``load``
manage tables.  Command line::

    python -m scalingelections_amd.autotune [--sizes 256,512,...] [--batch] [--out FILE]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time

import numpy as np

from . import _lib

WIDTHS = {"f16": _lib.SE_TUNE_F16, "u16": _lib.SE_TUNE_U16, "u32": _lib.SE_TUNE_U32}
WIDTH_FLAGS = {"f16": 0, "u16": _lib.SE_FLAG_NO_F16, "u32": _lib.SE_FLAG_FORCE_U32}
# (pivot, look-ahead mode): 0 off, 1 strips in a launch of their own, 2 one launch per
# round (16-bit widths, pivot 64)
CANDIDATES = ((32, 0), (32, 1), (64, 0), (64, 1), (64, 2))
LA_FLAGS = {0: _lib.SE_FLAG_NO_LOOKAHEAD, 1: _lib.SE_FLAG_LOOKAHEAD, 2: _lib.SE_FLAG_LOOKAHEAD_FUSED}
LA_NAMES = {0: "", 1: "+LA", 2: "+LA1"}


def candidates(width: str):
    return tuple(c for c in CANDIDATES if not (c[1] == 2 and width == "u32"))
DEFAULT_SIZES = (256, 512, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 12288, 16384)
BATCH_THRESHOLDS = (129, 193, 257, 385, 513, 1025)
ANY_SIDE = 0xFFFFFFFF
MACRO = 128  # the closure pads n up to a multiple of the 128-wide macro tile


def padded_side(n: int) -> int:
    return (int(n) + MACRO - 1) // MACRO * MACRO


# ------------------------------------------------------------------------ the table
def current_table() -> list[dict]:
    """The active closure table as dicts {width, max_side, pivot, lookahead}."""
    L = _lib.lib()
    buf = (_lib.se_tune_entry * _lib.SE_TUNE_MAX_ENTRIES)()
    count = ctypes.c_uint32(0)
    _lib.check(L.se_tune_closure_get(buf, _lib.SE_TUNE_MAX_ENTRIES, ctypes.byref(count)))
    names = {v: k for k, v in WIDTHS.items()}
    return [{"width": names[e.width], "max_side": int(e.max_side), "pivot": int(e.pivot),
             "lookahead": int(e.lookahead)} for e in buf[:count.value]]


def install(entries: list[dict] | None = None, batch_rounds_min: int | None = None) -> None:
    """Installs closure entries (None or [] restores the built-in table) and, when given,
    the batch threshold (0 restores the built-in value)."""
    L = _lib.lib()
    entries = entries or []
    arr = (_lib.se_tune_entry * max(1, len(entries)))()
    for i, e in enumerate(entries):
        arr[i] = _lib.se_tune_entry(WIDTHS[e["width"]], int(e["max_side"]), int(e["pivot"]), int(e["lookahead"]))
    _lib.check(L.se_tune_closure(arr, len(entries)))
    if batch_rounds_min is not None:
        _lib.check(L.se_tune_batch(int(batch_rounds_min)))


def batch_rounds_min() -> int:
    v = ctypes.c_uint32(0)
    _lib.check(_lib.lib().se_tune_batch_get(ctypes.byref(v)))
    return int(v.value)


def reset() -> None:
    install([], 0)


def save(path: str, entries: list[dict], batch_rounds_min_c: int | None = None, measurements=None) -> None:
    from .api import device_info
    doc = {"device": device_info(0), "closure": entries, "batch_rounds_min_candidates": batch_rounds_min_c,
           "measurements": measurements or {}}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


def load(path: str) -> dict:
    """Installs a table written by ``save``; returns the document."""
    with open(path) as f:
        doc = json.load(f)
    install(doc.get("closure") or [], doc.get("batch_rounds_min_candidates"))
    return doc


def buckets(best: dict[int, tuple[int, int]], width: str) -> list[dict]:
    """Table entries of one width from {padded side: (pivot, lookahead)}: each measured side
    bounds a bucket (sides between two measured ones take the larger one's choice), the
    largest extends to every side, and neighbours with the same choice merge."""
    sides = sorted(best)
    out: list[dict] = []
    for i, N in enumerate(sides):
        pivot, la = best[N]
        max_side = ANY_SIDE if i == len(sides) - 1 else N
        if out and out[-1]["pivot"] == pivot and out[-1]["lookahead"] == la:
            out[-1]["max_side"] = max_side
        else:
            out.append({"width": width, "max_side": max_side, "pivot": pivot, "lookahead": la})
    return out


# ------------------------------------------------------------------------ measuring
def _time_closure(L, dd, pp, n: int, flags: int, min_seconds: float, max_reps: int) -> float:
    from . import devmem
    best = float("inf")
    spent, reps = 0.0, 0
    while reps < 2 or (spent < min_seconds and reps < max_reps):
        devmem.device_synchronize()
        t0 = time.perf_counter()
        _lib.check(L.se_closure_device(dd.ptr, n, n, pp.ptr, n, flags, None, None))
        devmem.device_synchronize()
        el = time.perf_counter() - t0
        best = min(best, el)
        spent += el
        reps += 1
    return best


def tune_closure(sizes=DEFAULT_SIZES, widths=("f16", "u16", "u32"), min_seconds: float = 0.05,
                 max_reps: int = 50, verify_max_n: int = 8192, log=print):
    """Times every candidate (pivot, look-ahead) per width and size on synthesized
    preferences (seeded; values fit 16 bits with a span that allows the f16-ordered path,
    and the width is forced by flags).  Returns (entries, measurements)."""
    from . import devmem, synth
    L = _lib.lib()
    entries: list[dict] = []
    meas: dict = {}
    for width in widths:
        best: dict[int, tuple[int, int]] = {}
        for n in sizes:
            d = synth.synthesized_preferences(int(n), seed=3)
            dd = devmem.DeviceBuffer.from_array(d)
            pp = devmem.DeviceBuffer(int(n) * int(n) * 4)
            times = {}
            ref = None
            for pivot, la in candidates(width):
                flags = WIDTH_FLAGS[width] | (_lib.SE_FLAG_PIVOT32 if pivot == 32 else _lib.SE_FLAG_PIVOT64) \
                    | LA_FLAGS[la]
                times[(pivot, la)] = _time_closure(L, dd, pp, int(n), flags, min_seconds, max_reps)
                if n <= verify_max_n:
                    out = pp.download(np.empty((int(n), int(n)), np.uint32))
                    if ref is None:
                        ref = out
                    elif not np.array_equal(ref, out):
                        raise RuntimeError(f"closure result differs between launch shapes at n={n} {width} "
                                           f"pivot={pivot} lookahead={la}")
            choice = min(times, key=times.get)
            best[padded_side(n)] = choice
            meas[f"{width}/{n}"] = {f"T{p}{LA_NAMES[a]}": round(t * 1e3, 4) for (p, a), t in times.items()}
            if log:
                log(f"{width:4s} n={n:6d}: " + "  ".join(f"T{p}{LA_NAMES[a]:4s} {t * 1e3:9.3f} ms"
                                                          for (p, a), t in sorted(times.items()))
                    + f"  -> T{choice[0]}{LA_NAMES[choice[1]]}")
            dd.close()
            pp.close()
        entries += buckets(best, width)
    return entries, meas


def tune_batch(elections: int = 4096, seed: int = 4, thresholds=BATCH_THRESHOLDS, reps: int = 5, log=print):
    """Times the batch's closure launches (HIP events; the threshold does not touch the
    counting) on the mixed-size workload (C log-uniform in [8, 512], m uniform in
    [100, 10000]) for every threshold; returns (best threshold, {threshold: ms})."""
    from . import batch, devmem
    w = batch.mixed_size_workload(elections, seed=seed)
    db = batch.DeviceBatch(w)
    before = batch_rounds_min()
    times, ref = {}, None
    try:
        for th in thresholds:
            _lib.check(_lib.lib().se_tune_batch(int(th)))
            db.run()
            devmem.device_synchronize()
            best = float("inf")
            for _ in range(reps):
                best = min(best, db.run()["closure_ms"] / 1e3)
            out = db.raw()
            if ref is None:
                ref = out
            elif not (np.array_equal(ref[0], out[0]) and np.array_equal(ref[1], out[1])):
                raise RuntimeError(f"batch result differs at threshold {th}")
            times[int(th)] = round(best * 1e3, 4)
            if log:
                log(f"batch rounds_min={th:5d}: closure {best * 1e3:8.3f} ms")
    finally:
        _lib.check(_lib.lib().se_tune_batch(before))
    return min(times, key=times.get), times


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Measure and install the fastest launch shapes on this GPU.")
    ap.add_argument("--sizes", default=",".join(str(s) for s in DEFAULT_SIZES))
    ap.add_argument("--widths", default="f16,u16,u32")
    ap.add_argument("--batch", action="store_true", help="also tune the batch threshold")
    ap.add_argument("--out", default=None, help="write the table as JSON (load with autotune.load)")
    args = ap.parse_args(argv)
    sizes = [int(s) for s in args.sizes.split(",") if s]
    entries, meas = tune_closure(sizes, tuple(args.widths.split(",")))
    th = None
    if args.batch:
        th, bmeas = tune_batch()
        meas["batch"] = bmeas
    install(entries, th)
    print(json.dumps({"closure": entries, "batch_rounds_min_candidates": th}, indent=1))
    if args.out:
        save(args.out, entries, th, meas)
    return 0


if __name__ == "__main__":
    sys.exit(main())
