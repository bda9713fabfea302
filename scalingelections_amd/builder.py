"""Preference-matrix builder: ballots -> positions -> the GPU counter.

Reference: ``build_pairwise_preferences`` and ``populate_preferences_from_ranking``,
scaling_elections.py:37-88.  The rules kept exactly:
  * n = max(1, max id + 1) over all ballots (:63-66);
  * a ballot shorter than n gets the ids it lacks appended in ascending order (:74-83);
  * every pair ranked i before j adds one to d[i][j], uint32 (:46-48).
Ballots without repeated ids become position arrays (positions[v][c] = place of c) and
are counted by the dense compare-and-count kernel.  Ballots that repeat an id -- which the
reference also counts, touching the diagonal -- go to the scatter kernel that replays the
reference's pair loop exactly.  Either way the counting runs on the GPU; there is no CPU
path.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import _lib


def count_candidates(voter_rankings: Sequence[np.ndarray]) -> int:
    """n = max(1, max id + 1), scaling_elections.py:64-66 (np.max raises on an empty ballot,
    as in the reference)."""
    n = 1
    for ranking in voter_rankings:
        n = max(n, int(np.max(ranking)) + 1)
    return n


def _ids(ranking) -> np.ndarray:
    r = np.asarray(ranking)
    if r.dtype.kind not in "iu":
        r = r.astype(np.int64)
    r = r.ravel()
    if r.size and int(r.min()) < 0:
        raise ValueError("candidate ids must be non-negative")
    return r


def rankings_to_positions(voter_rankings: Sequence[np.ndarray], n: int):
    """positions (m x n int32) for ballots without repeated ids, filling incomplete
    ballots in ascending id order; returns None when some ballot repeats an id."""
    m = len(voter_rankings)
    ar = np.arange(n, dtype=np.int32)
    # fast path: every ballot complete and of equal length -> one scatter
    if m and all(np.ndim(r) == 1 and len(r) == n for r in voter_rankings):
        R = np.asarray(voter_rankings)
        if R.dtype.kind in "iu" and R.shape == (m, n) and (n == 0 or int(R.min()) >= 0):
            pos = np.full((m, n), -1, dtype=np.int32)
            np.put_along_axis(pos, R.astype(np.intp), np.broadcast_to(ar, (m, n)), axis=1)
            if (pos >= 0).all():
                return pos
            return None
    pos = np.empty((m, n), dtype=np.int32)
    for v, ranking in enumerate(voter_rankings):
        r = _ids(ranking)
        mask = np.ones(n, dtype=bool)
        mask[r] = False
        if r.size + int(mask.sum()) != n:  # an id appears twice
            return None
        row = np.empty(n, dtype=np.int32)
        row[r] = ar[: r.size]
        missing = np.nonzero(mask)[0]
        row[missing] = ar[r.size:]
        pos[v] = row
    return pos


def filled_ballots(voter_rankings: Sequence[np.ndarray], n: int):
    """Concatenated ballots after the ascending fill (scaling_elections.py:75-83), plus
    offsets -- the input of the exact scatter counter."""
    parts = []
    offsets = np.zeros(len(voter_rankings) + 1, dtype=np.int64)
    for v, ranking in enumerate(voter_rankings):
        r = _ids(ranking)
        if r.size != n:
            mask = np.ones(n, dtype=bool)
            mask[r] = False
            r = np.append(r, np.nonzero(mask)[0])
        parts.append(r.astype(np.int32))
        offsets[v + 1] = offsets[v] + r.size
    ids = np.concatenate(parts) if parts else np.zeros(0, dtype=np.int32)
    return np.ascontiguousarray(ids, dtype=np.int32), offsets


def count_positions(positions: np.ndarray) -> np.ndarray:
    """d (n x n uint32) from an m x n int32 position matrix whose rows are strict orders."""
    pos = np.ascontiguousarray(positions, dtype=np.int32)
    if pos.ndim != 2:
        raise ValueError("positions must be an (m, n) array")
    m, n = pos.shape
    d = np.empty((n, n), dtype=np.uint32)
    if n == 0:
        return d
    _lib.check(_lib.lib().se_count(_lib.ptr(pos) if m else None, m, n, _lib.ptr(d), n))
    return d


def build_pairwise_preferences(voter_rankings: Sequence[np.ndarray]) -> np.ndarray:
    """d[i][j] = number of voters ranking i above j (scaling_elections.py:51-88)."""
    n = count_candidates(voter_rankings)
    pos = rankings_to_positions(voter_rankings, n)
    if pos is not None:
        return count_positions(pos)
    ids, offsets = filled_ballots(voter_rankings, n)
    d = np.empty((n, n), dtype=np.uint32)
    _lib.check(_lib.lib().se_count_rankings(_lib.ptr(ids) if ids.size else None, _lib.ptr(offsets),
                                            len(voter_rankings), n, _lib.ptr(d), n))
    return d
