"""Shared helpers for the test suite: fixtures, independent brute-force references."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name: str) -> dict:
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def schulze45_ballots():
    g = load_golden("schulze_45.json")
    ballots = []
    for count, order in g["groups"]:
        ballots += [np.array(["ABCDE".index(c) for c in order])] * count
    return ballots, g


def p0_numpy(d: np.ndarray) -> np.ndarray:
    """p0 by the rule p0[i][j] = d[i][j] if d[i][j] > d[j][i] else 0, diagonal 0 (vectorised,
    written independently of the oracle's loops)."""
    d = np.asarray(d, dtype=np.uint64)
    n = d.shape[0]
    p0 = np.where(d > d.T, d, 0)
    p0[np.arange(n), np.arange(n)] = 0
    return p0.astype(np.uint64)


def brute_widest(d: np.ndarray) -> np.ndarray:
    """Widest-path closure by per-threshold reachability -- no Floyd-Warshall involved.
    p[i][j] = max t such that j is reachable from i using only edges with p0 >= t."""
    p0 = p0_numpy(d)
    n = p0.shape[0]
    out = np.zeros((n, n), dtype=np.uint64)
    for t in sorted(set(p0.ravel().tolist()) - {0}):
        R = p0 >= t
        while True:  # boolean transitive closure by repeated squaring
            R2 = R | ((R.astype(np.int64) @ R.astype(np.int64)) > 0)
            if (R2 == R).all():
                break
            R = R2
        out[R] = t
    out[np.arange(n), np.arange(n)] = 0
    return out.astype(np.uint32)


def brute_count(ballots, n: int) -> np.ndarray:
    """prefs[r[i]][r[j]] += 1 for i < j over ballots already filled to full length."""
    d = np.zeros((n, n), dtype=np.uint64)
    for r in ballots:
        r = [int(x) for x in r]
        for i in range(len(r)):
            for j in range(i + 1, len(r)):
                d[r[i], r[j]] += 1
    return d.astype(np.uint32)


def fill_ballot(r, n: int):
    """Ascending fill of missing ids (scaling_elections.py:75-83), written as a plain loop."""
    r = [int(x) for x in r]
    if len(r) == n:
        return r
    present = set(r)
    return r + [c for c in range(n) if c not in present]


def python_ranking(wins) -> list:
    return sorted(range(len(wins)), key=lambda x: wins[x], reverse=True)


def closure_is_closed(p: np.ndarray) -> bool:
    """p[i][j] >= min(p[i][k], p[k][j]) for all distinct i, j, k."""
    p = np.asarray(p, dtype=np.int64)
    n = p.shape[0]
    for k in range(n):
        m = np.minimum(p[:, k:k + 1], p[k:k + 1, :])
        m[k, :] = 0
        m[:, k] = 0
        np.fill_diagonal(m, 0)
        if (m > p).any():
            return False
    return True
