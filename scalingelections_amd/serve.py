"""Continuous-batching election server.

Requests arrive over time (Poisson arrivals in the benchmark) and ask for one election's
strongest paths and ranking.  The server:

  * admits requests in arrival order (per size class) while the device-memory footprint of
    everything in flight stays within a budget taken from free device memory at start;
  * runs small elections (<= 1024 candidates) in batches: every pass of the loop packs all
    admitted small requests into one ``se_batch_device`` call on the batch stream; requests
    that arrive while a batch runs join the next one;
  * runs large elections from worker threads (one per large slot, each with its own
    stream) through the single-election entry points (counter, closure, ranking), so they
    overlap with the batches and with each other on the device;
  * keeps inputs and results in device memory (preallocated arenas and slots, no per-request
    hipMalloc) and copies back the ranking, plus the strongest-path matrix when asked.

Every request's result is, bit for bit, the one the election gets when processed alone.
"""
from __future__ import annotations

import ctypes
import hashlib
import queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib, devmem
from .batch import Workload

KMAX_BATCH_C = 1024


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the run starts
    pool: Workload            # where its ballots live
    index: int                # election index in the pool
    admitted: float = -1.0
    done: float = -1.0
    ranking: Optional[np.ndarray] = None
    paths_sha256: Optional[str] = None

    @property
    def C(self) -> int:
        return int(self.pool.C[self.index])

    @property
    def m(self) -> int:
        return int(self.pool.m[self.index])

    @property
    def large(self) -> bool:
        return self.C > KMAX_BATCH_C

    def footprint(self) -> int:
        """Device bytes while in flight: positions, d, p, work matrix, ranking."""
        C, m = self.C, self.m
        cp = ((C + 127) // 128) * 128
        return m * C * 4 + 2 * C * C * 4 + cp * cp * 4 + C * 4


def poisson_schedule(rate: float, count: int, seed: int) -> np.ndarray:
    """Arrival times (s) of `count` requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


try:  # 128-bit xxh3 digests strongest-path matrices at memory speed; sha256 otherwise
    import xxhash as _xxhash
except ImportError:  # pragma: no cover - xxhash ships with this environment
    _xxhash = None


def _sha(a: np.ndarray) -> str:
    """Digest of a strongest-path matrix (little-endian uint32 bytes)."""
    buf = np.ascontiguousarray(a, dtype="<u4")
    if _xxhash is not None:
        return "xxh3_128:" + _xxhash.xxh3_128_hexdigest(memoryview(buf).cast("B"))
    return "sha256:" + hashlib.sha256(buf.tobytes()).hexdigest()


@dataclass
class ServeResult:
    requests: List[Request]
    wall_s: float
    batches: int
    batch_sizes: List[int] = field(default_factory=list)

    def latencies(self, cls: Optional[str] = None) -> np.ndarray:
        rs = [r for r in self.requests if cls is None or (cls == "large") == r.large]
        return np.array([r.done - r.arrival for r in rs], dtype=np.float64)

    def summary(self) -> dict:
        out = {"requests": len(self.requests), "wall_s": round(self.wall_s, 3),
               "completed_per_s": round(len(self.requests) / self.wall_s, 1), "batches": self.batches,
               "mean_batch": round(float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0, 2)}
        for cls in (None, "small", "large"):
            lat = self.latencies(cls)
            key = cls or "all"
            if lat.size:
                out[key] = {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                            "max_ms": round(float(lat.max()) * 1e3, 3)}
        return out


class Server:
    def __init__(self, budget_bytes: Optional[int] = None, small_arena_bytes: int = 1 << 30,
                 large_slots: int = 2, large_max_c: int = 8192, large_max_m: int = 8192,
                 keep_paths: bool = True):
        free, _ = devmem.mem_info()
        self.budget = int(min(budget_bytes or free * 0.6, free * 0.8))
        self.keep_paths = keep_paths
        self.s_small = devmem.Stream(nonblocking=True)
        self.s_large = [devmem.Stream(nonblocking=True) for _ in range(max(1, large_slots))]
        # small-batch arena: positions | paths | ranking
        self.arena_pos = devmem.DeviceBuffer(small_arena_bytes)
        self.arena_paths = devmem.DeviceBuffer(small_arena_bytes // 2)
        self.arena_rank = devmem.DeviceBuffer(16 << 20)
        # large-request slots
        self.large_cap = (large_max_c, large_max_m)
        self.slots = queue.Queue()
        for i in range(max(1, large_slots)):
            self.slots.put({
                "pos": devmem.DeviceBuffer(large_max_c * large_max_m * 4),
                "d": devmem.DeviceBuffer(large_max_c * large_max_c * 4),
                "p": devmem.DeviceBuffer(large_max_c * large_max_c * 4),
                "stream": self.s_large[i],
            })
        self._lock = threading.Lock()
        self._inflight = 0

    def close(self):
        for s in [self.s_small] + self.s_large:
            s.close()

    # ------------------------------------------------------------------ accounting
    def _try_reserve(self, nbytes: int) -> bool:
        with self._lock:
            if self._inflight + nbytes > self.budget:
                return False
            self._inflight += nbytes
            return True

    def _release(self, nbytes: int) -> None:
        with self._lock:
            self._inflight -= nbytes

    # ------------------------------------------------------------------ large path
    def _run_large(self, r: Request, t0: float, slot: dict) -> None:
        L = _lib.lib()
        C, m = r.C, r.m
        pos = r.pool.election_positions(r.index)
        stream = slot["stream"]
        slot["pos"].upload_async(pos, stream)
        sh = ctypes.c_void_p(stream.handle)
        _lib.check(L.se_count_device(slot["pos"].ptr, m, C, slot["d"].ptr, C, sh, None))
        _lib.check(L.se_closure_device(slot["d"].ptr, C, C, slot["p"].ptr, C, 0, sh, None))
        wins = np.empty(C, np.uint32)
        ranking = np.empty(C, np.uint32)
        _lib.check(L.se_rank_device(slot["p"].ptr, C, C, wins.ctypes.data, ranking.ctypes.data, sh))
        r.done = time.perf_counter() - t0
        r.ranking = ranking
        if self.keep_paths:
            p = np.empty((C, C), np.uint32)
            slot["p"].download_async(p, stream)
            stream.synchronize()
            r.paths_sha256 = _sha(p)

    def _large_worker(self, q: "queue.Queue", t0: float):
        while True:
            item = q.get()
            if item is None:
                return
            r, slot = item
            try:
                self._run_large(r, t0, slot)
            finally:
                self.slots.put(slot)
                self._release(r.footprint())

    # ------------------------------------------------------------------ small path
    def _run_batch(self, batch: List[Request], t0: float) -> None:
        C = np.array([r.C for r in batch], np.uint32)
        m = np.array([r.m for r in batch], np.uint32)
        parts = [r.pool.election_positions(r.index).ravel() for r in batch]
        pos = np.concatenate(parts) if parts else np.zeros(1, np.int32)
        self.arena_pos.upload_async(pos, self.s_small)
        sh = ctypes.c_void_p(self.s_small.handle)
        _lib.check(_lib.lib().se_batch_device(len(batch), C.ctypes.data, m.ctypes.data, self.arena_pos.ptr,
                                              self.arena_paths.ptr, self.arena_rank.ptr, 0, sh, None))
        rtot = int(C.astype(np.int64).sum())
        rank = np.empty(max(1, rtot), np.uint32)
        self.arena_rank.download_async(rank[:rtot], self.s_small)
        self.s_small.synchronize()
        now = time.perf_counter() - t0
        ro = 0
        for r in batch:
            r.done = now
            r.ranking = rank[ro:ro + r.C].copy()
            ro += r.C
        if self.keep_paths:
            ptot = int((C.astype(np.int64) ** 2).sum())
            paths = np.empty(max(1, ptot), np.uint32)
            self.arena_paths.download_async(paths[:ptot], self.s_small)
            self.s_small.synchronize()
            po = 0
            for r in batch:
                r.paths_sha256 = _sha(paths[po:po + r.C * r.C])
                po += r.C * r.C
        del pos

    def _batch_fits(self, C, m, pos_bytes, paths_elems, rank_elems) -> bool:
        return (pos_bytes <= self.arena_pos.nbytes and paths_elems * 4 <= self.arena_paths.nbytes
                and rank_elems * 4 <= self.arena_rank.nbytes)

    # ------------------------------------------------------------------ loop
    def run(self, requests: List[Request], idle_sleep: float = 2e-4) -> ServeResult:
        for r in requests:
            if r.large and (r.C > self.large_cap[0] or r.m > self.large_cap[1]):
                raise ValueError("request exceeds the large-request slot size")
            if r.footprint() > self.budget:
                raise ValueError("request larger than the whole device-memory budget")
        reqs = sorted(requests, key=lambda r: r.arrival)
        N = len(reqs)
        lq: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()
        workers = [threading.Thread(target=self._large_worker, args=(lq, t0), daemon=True)
                   for _ in range(len(self.s_large))]
        for wk in workers:
            wk.start()
        small, large = deque(), deque()
        i = 0
        batches, sizes = 0, []
        try:
            while True:
                now = time.perf_counter() - t0
                while i < N and reqs[i].arrival <= now:
                    (large if reqs[i].large else small).append(reqs[i])
                    i += 1
                # large: admitted when a slot and memory are free
                while large and not self.slots.empty():
                    r = large[0]
                    if not self._try_reserve(r.footprint()):
                        break
                    slot = self.slots.get()
                    r.admitted = now
                    lq.put((r, slot))
                    large.popleft()
                # small: everything that fits the memory budget and the batch arena
                batch, pos_b, pe, re_ = [], 0, 0, 0
                while small:
                    r = small[0]
                    nb, npth, nr = r.m * r.C * 4, r.C * r.C, r.C
                    if not self._batch_fits(None, None, pos_b + nb, pe + npth, re_ + nr):
                        break
                    if not self._try_reserve(r.footprint()):
                        break
                    r.admitted = now
                    batch.append(r)
                    small.popleft()
                    pos_b, pe, re_ = pos_b + nb, pe + npth, re_ + nr
                if batch:
                    self._run_batch(batch, t0)
                    batches += 1
                    sizes.append(len(batch))
                    for r in batch:
                        self._release(r.footprint())
                    continue
                if i >= N and not small and not large:
                    break
                nxt = reqs[i].arrival - (time.perf_counter() - t0) if i < N else idle_sleep
                time.sleep(max(0.0, min(idle_sleep, nxt)))
        finally:
            for _ in workers:
                lq.put(None)
            for wk in workers:
                wk.join()
        wall = max(r.done for r in reqs) if reqs else 0.0
        return ServeResult(reqs, wall, batches, sizes)


def solo_result(pool: Workload, index: int, stream: Optional[devmem.Stream] = None, timed: bool = False):
    """The election processed alone through the single-election entry points:
    (ranking, digest of the strongest-path matrix[, seconds of the whole service the server
    gives a large request: upload, count, closure, ranking, download, digest])."""
    L = _lib.lib()
    C, m = int(pool.C[index]), int(pool.m[index])
    host_pos = pool.election_positions(index)
    pos = devmem.DeviceBuffer(max(4, host_pos.nbytes))
    d = devmem.DeviceBuffer(C * C * 4)
    p = devmem.DeviceBuffer(C * C * 4)
    sh = ctypes.c_void_p(stream.handle) if stream else None
    devmem.device_synchronize()
    t0 = time.perf_counter()
    if host_pos.size:
        pos.upload(host_pos)
    _lib.check(L.se_count_device(pos.ptr, m, C, d.ptr, C, sh, None))
    _lib.check(L.se_closure_device(d.ptr, C, C, p.ptr, C, 0, sh, None))
    wins = np.empty(C, np.uint32)
    ranking = np.empty(C, np.uint32)
    _lib.check(L.se_rank_device(p.ptr, C, C, wins.ctypes.data, ranking.ctypes.data, sh))
    out = np.empty((C, C), np.uint32)
    devmem.device_synchronize()
    p.download(out)
    digest = _sha(out)
    elapsed = time.perf_counter() - t0
    if timed:
        return ranking, digest, elapsed
    return ranking, digest
