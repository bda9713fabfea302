"""Batches of independent elections of uneven size: one device call for the whole batch.

Each election e has C_e candidates and m_e ballots (positions m_e x C_e, strict orders).
The batch returns, per election, the strongest-path matrix and the ranking that
``build_pairwise_preferences`` + ``compute_strongest_paths`` + ``get_winner_and_ranking``
(scaling_elections.py:51-296, scaling_elections.cu:863-989) return for that election alone.

Strategies (same outputs):
  * batched (default): one counting launch over work items for the whole batch, then one
    workgroup per election for p0, closure, wins and ranking -- elections above 128
    candidates instead share one pair of launches per round (panels, then tiles);
  * per-launch: every election through the single-election entry points, one after the
    other -- the baseline the batched form is measured against.

Workload generator ``mixed_size_workload``: C log-uniform in [8, 512], m uniform in
[100, 10000], Mallows ballots with dispersion phi uniform in [0.5, 1.0] around a seeded
reference order (seed 4 for the benchmark).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib, devmem


@dataclass
class Workload:
    C: np.ndarray          # uint32 [E]
    m: np.ndarray          # uint32 [E]
    phi: np.ndarray        # float64 [E]
    reference: np.ndarray  # int32, concatenated reference orders
    positions: np.ndarray  # int32, concatenated m_e x C_e position arrays
    seed: int

    @property
    def elections(self) -> int:
        return int(self.C.shape[0])

    def pos_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.C.astype(np.int64) * self.m.astype(np.int64))])

    def paths_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.C.astype(np.int64) ** 2)])

    def rank_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.C.astype(np.int64))])

    def election_positions(self, e: int) -> np.ndarray:
        o = self.pos_offsets()
        return self.positions[o[e]:o[e + 1]].reshape(int(self.m[e]), int(self.C[e]))


def mallows_positions(C, m, phi, reference, seed: int, threads: int = 0) -> np.ndarray:
    C = np.ascontiguousarray(C, dtype=np.uint32)
    m = np.ascontiguousarray(m, dtype=np.uint32)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    reference = np.ascontiguousarray(reference, dtype=np.int32)
    total = int((C.astype(np.int64) * m.astype(np.int64)).sum())
    out = np.empty(max(1, total), dtype=np.int32)
    _lib.check(_lib.lib().se_synth_mallows(C.shape[0], _lib.ptr(C), _lib.ptr(m), _lib.ptr(phi),
                                           _lib.ptr(reference) if reference.size else None, seed,
                                           _lib.ptr(out), threads))
    return out[:total]


def make_workload(C, m, phi, seed: int) -> Workload:
    rng = np.random.default_rng([seed, 1])
    C = np.asarray(C, dtype=np.uint32)
    refs = [rng.permutation(int(c)).astype(np.int32) for c in C]
    reference = np.concatenate(refs) if refs else np.zeros(0, np.int32)
    positions = mallows_positions(C, m, phi, reference, seed)
    return Workload(C, np.asarray(m, dtype=np.uint32), np.asarray(phi, dtype=np.float64), reference, positions, seed)


def mixed_size_workload(elections: int = 4096, seed: int = 4, c_range=(8, 512), m_range=(100, 10000),
                        phi_range=(0.5, 1.0)) -> Workload:
    rng = np.random.default_rng(seed)
    C = np.clip(np.round(np.exp(rng.uniform(np.log(c_range[0]), np.log(c_range[1]), elections))),
                c_range[0], c_range[1]).astype(np.uint32)
    m = rng.integers(m_range[0], m_range[1] + 1, elections).astype(np.uint32)
    phi = rng.uniform(phi_range[0], phi_range[1], elections)
    return make_workload(C, m, phi, seed)


class DeviceBatch:
    """A workload uploaded to the device with output buffers, ready to run repeatedly.

    pos16 (default): the positions are kept on the device as uint16 (converted there by
    se_positions16_device) and the batch runs with SE_BATCH_POS16, halving the bytes the
    counter reads; False keeps the int32 positions.  Results are the same either way."""

    def __init__(self, w: Workload, pos16: bool = True):
        self.w = w
        self.pos16 = bool(pos16) and w.positions.size > 0
        if not w.positions.size:
            self.pos = devmem.DeviceBuffer(4)
        elif self.pos16:
            raw = devmem.DeviceBuffer.from_array(w.positions)
            self.pos = devmem.DeviceBuffer(w.positions.size * 2)
            _lib.check(_lib.lib().se_positions16_device(raw.ptr, w.positions.size, self.pos.ptr, None))
            raw.close()
        else:
            self.pos = devmem.DeviceBuffer.from_array(w.positions)
        self.paths = devmem.DeviceBuffer(int(w.paths_offsets()[-1]) * 4 + 4)
        self.ranking = devmem.DeviceBuffer(int(w.rank_offsets()[-1]) * 4 + 4)
        self.C = np.ascontiguousarray(w.C, dtype=np.uint32)
        self.m = np.ascontiguousarray(w.m, dtype=np.uint32)

    def run(self, per_launch: bool = False, stream=None, one_group: bool = False) -> dict:
        t = _lib.se_batch_timing()
        flags = (_lib.SE_BATCH_PER_LAUNCH if per_launch else 0) | (_lib.SE_BATCH_ONE_GROUP if one_group else 0) \
            | (_lib.SE_BATCH_POS16 if self.pos16 else 0)
        rc = _lib.lib().se_batch_device(self.C.shape[0], _lib.ptr(self.C), _lib.ptr(self.m), self.pos.ptr,
                                        self.paths.ptr, self.ranking.ptr, flags,
                                        ctypes.c_void_p(stream.handle) if stream else None, ctypes.byref(t))
        _lib.check(rc)
        return t.as_dict()

    def results(self):
        """(paths, rankings): lists of per-election host arrays."""
        po, ro = self.w.paths_offsets(), self.w.rank_offsets()
        P = self.paths.download(np.empty(int(po[-1]), dtype=np.uint32))
        R = self.ranking.download(np.empty(int(ro[-1]), dtype=np.uint32))
        paths = [P[po[e]:po[e + 1]].reshape(int(self.C[e]), int(self.C[e])) for e in range(self.C.shape[0])]
        ranks = [R[ro[e]:ro[e + 1]] for e in range(self.C.shape[0])]
        return paths, ranks

    def raw(self):
        po, ro = self.w.paths_offsets(), self.w.rank_offsets()
        return (self.paths.download(np.empty(int(po[-1]), dtype=np.uint32)),
                self.ranking.download(np.empty(int(ro[-1]), dtype=np.uint32)))
