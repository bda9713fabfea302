"""Winner and ranking from a strongest-path matrix.

Reference: ``get_winner_and_ranking``, scaling_elections.py:274-296.  Win counts run on the
GPU (wins[i] = #{j != i : p[i][j] > p[j][i]}); the ranking is a stable descending sort by
wins, ties in ascending index -- Python's ``sorted(..., reverse=True)`` keeps equal keys in
their original order (:292).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import _lib


def _paths_u32(strongest_paths, k: int) -> np.ndarray:
    p = np.asarray(strongest_paths)
    if p.ndim != 2 or p.shape[0] < k or p.shape[1] < k:
        raise IndexError("strongest_paths is smaller than the candidate list")
    p = p[:k, :k]
    if p.dtype != np.uint32:
        if p.dtype.kind not in "iub":
            raise TypeError("strongest_paths must hold integers")
        if p.size and (int(p.min()) < 0 or int(p.max()) > 0xFFFFFFFF):
            raise ValueError("strongest_paths values must fit in uint32")
        p = p.astype(np.uint32)
    return np.ascontiguousarray(p)


def wins_and_order(strongest_paths, k: int = None) -> Tuple[np.ndarray, np.ndarray]:
    """(wins, ranking indices) for the leading k x k block (default: the whole matrix)."""
    if k is None:
        k = np.asarray(strongest_paths).shape[0]
    p = _paths_u32(strongest_paths, k)
    wins = np.zeros(k, dtype=np.uint32)
    order = np.zeros(k, dtype=np.uint32)
    if k:
        _lib.check(_lib.lib().se_rank(_lib.ptr(p), k, k, _lib.ptr(wins), _lib.ptr(order)))
    return wins, order


def get_winner_and_ranking(candidates: list, strongest_paths: np.ndarray) -> Tuple[int, List[int]]:
    """(winner, ranked candidates), scaling_elections.py:274-296."""
    num_candidates = len(candidates)
    _, order = wins_and_order(strongest_paths, num_candidates)
    ranking_indices = [int(i) for i in order]
    winner = candidates[ranking_indices[0]]  # IndexError for no candidates, as in the reference
    ranked_candidates = [candidates[i] for i in ranking_indices]
    return winner, ranked_candidates
