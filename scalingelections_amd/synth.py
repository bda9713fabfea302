"""Seeded synthetic elections for the benchmark and parity workloads.

Every generator is deterministic given its seed (numpy ``default_rng`` streams keyed by
(seed, block index)), independent of the number of worker threads, so the CPU oracle and
the GPU run see identical inputs on any machine with the same numpy.

Workloads:
  * ``impartial_culture(m, n, seed)``: m ballots, each a uniformly random permutation of
    0..n-1 (128 x 2000 with seed 1; 4096 x 4096 with seed 2);
  * ``synthesized_preferences(n, seed)``: the roofline-scale closure input -- for i < j,
    d[i][j] ~ Binomial(65535, clip(0.5 + 0.02 z, 0, 1)) with z ~ N(0, 1) and
    d[j][i] = 65535 - d[i][j]; a Bernoulli(0.05) mask forces exact ties, both cells 32767
    (65535 is odd, so an exact split is impossible); diagonal 0 (n = 32768, seed 3);
  * ``mallows_election(...)``: Mallows ballots by repeated insertion around a reference
    order (mixed-size batches).
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np

VOTES_C3 = 65535
TIE_VALUE = 32767
_FLAG = np.uint32(1 << 31)


def _workers() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def impartial_culture(m: int, n: int, seed: int) -> np.ndarray:
    """(m, n) int32 rankings: row v lists candidate ids in voter v's preference order."""
    rng = np.random.default_rng(seed)
    base = np.broadcast_to(np.arange(n, dtype=np.int32), (m, n))
    return np.ascontiguousarray(rng.permuted(base, axis=1), dtype=np.int32)


def positions_from_rankings(rankings: np.ndarray) -> np.ndarray:
    """Inverse permutations: positions[v][c] = place of candidate c in ballot v."""
    R = np.asarray(rankings)
    m, n = R.shape
    pos = np.empty((m, n), dtype=np.int32)
    np.put_along_axis(pos, R.astype(np.intp), np.broadcast_to(np.arange(n, dtype=np.int32), (m, n)), axis=1)
    return pos


def _c3_block(d: np.ndarray, seed: int, b: int, r0: int, r1: int, q_sd: float, tie_fraction: float):
    n = d.shape[0]
    rng = np.random.default_rng([seed, b])
    for i in range(r0, r1):
        cnt = n - i - 1
        if cnt <= 0:
            continue
        z = rng.standard_normal(cnt)
        q = np.clip(0.5 + q_sd * z, 0.0, 1.0)
        v = rng.binomial(VOTES_C3, q).astype(np.uint32)
        tie = rng.random(cnt) < tie_fraction
        v[tie] = np.uint32(TIE_VALUE) | _FLAG
        d[i, i + 1:] = v


def _c3_mirror(d: np.ndarray, I0: int, I1: int, tile: int):
    """Lower triangle of rows [I0, I1) from the flagged upper triangle, tile by tile:
    d[i][j] = 32767 where the pair is a forced tie, else 65535 - d[j][i] (j < i).  Reads only
    upper cells and writes only lower cells, so row blocks run concurrently."""
    for J0 in range(0, I1, tile):
        J1 = min(J0 + tile, I1)
        up = d[J0:J1, I0:I1].T  # d[j][i] for j in [J0, J1), i in [I0, I1)
        tie = (up & _FLAG) != 0
        lower = np.where(tie, np.uint32(TIE_VALUE), np.uint32(VOTES_C3) - (up & np.uint32(0xFFFF)))
        blk = d[I0:I1, J0:J1]
        ii = np.arange(I0, I1)[:, None]
        jj = np.arange(J0, J1)[None, :]
        sel = jj < ii
        blk[sel] = lower[sel]


def synthesized_preferences(n: int, seed: int = 3, q_sd: float = 0.02, tie_fraction: float = 0.05,
                            workers: int = 0) -> np.ndarray:
    """The roofline-scale preference matrix (uint32, n x n), see the module docstring."""
    d = np.zeros((n, n), dtype=np.uint32)
    rows = 64
    blocks = [(b, r0, min(n, r0 + rows)) for b, r0 in enumerate(range(0, n, rows))]
    W = workers or _workers()
    with cf.ThreadPoolExecutor(W) as ex:
        list(ex.map(lambda a: _c3_block(d, seed, a[0], a[1], a[2], q_sd, tie_fraction), blocks))
    tile = 512
    with cf.ThreadPoolExecutor(W) as ex:
        list(ex.map(lambda r0: _c3_mirror(d, r0, min(n, r0 + tile), tile), range(0, n, tile)))
    np.bitwise_and(d, np.uint32(0xFFFF), out=d)
    return d


def mallows_rankings(m: int, n: int, phi: float, rng: np.random.Generator,
                     reference: np.ndarray = None) -> np.ndarray:
    """(m, n) int32 Mallows ballots by repeated insertion: item i of the reference order is
    inserted at position j <= i with probability proportional to phi^(i - j)
    (phi = 1 is impartial culture)."""
    if reference is None:
        reference = rng.permutation(n).astype(np.int32)
    out = np.empty((m, n), dtype=np.int32)
    # insertion offsets from the end: P(off = t) ~ phi^t, t in [0, i]
    u = rng.random((m, n))
    i = np.arange(n, dtype=np.float64)
    if phi >= 1.0:
        off = np.floor(u * (i + 1)).astype(np.int64)
    else:
        # inverse CDF of the truncated geometric on {0..i}
        lp = np.log(phi)
        tot = 1.0 - np.power(phi, i + 1)
        off = np.floor(np.log1p(-u * tot) / lp).astype(np.int64)
    off = np.minimum(off, np.arange(n)[None, :])
    for v in range(m):
        ins = []
        for k in range(n):
            ins.insert(k - int(off[v, k]), int(reference[k]))
        out[v] = ins
    return out
