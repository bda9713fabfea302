"""Writes the 32768-candidate parity fixture: sha256 of the synthesized input d, of the
oracle's strongest-path matrix p (uint32 little-endian, row-major) and of its ranking.

The closure is the oracle's tiled restatement of the reference's Numba tiled kernel
(tile size and padding independence are pinned in tests/test_oracle.py).  Runs for tens of
minutes on 8 cores; rerun only when the generator or the seed changes:
    python tests/golden/make_c3_fixture.py [n] [seed]
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from scalingelections_amd import synth  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).astype("<u4", copy=False).tobytes()).hexdigest()


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 32768
    seed = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    t0 = time.time()
    d = synth.synthesized_preferences(n, seed)
    t1 = time.time()
    print(f"generated d in {t1 - t0:.1f}s", flush=True)
    p = oracle.closure_tiled(d, 64, threads=os.cpu_count() or 1)
    t2 = time.time()
    print(f"oracle closure in {t2 - t1:.1f}s", flush=True)
    order = oracle.ranking(p) if n <= 4096 else None
    wins = None
    if order is None:
        # O(n^2) wins in numpy blocks, then the stable descending sort of
        # scaling_elections.py:292 via a stable argsort on -wins
        wins = np.zeros(n, dtype=np.int64)
        B = 2048
        for i0 in range(0, n, B):
            blk = p[i0:i0 + B]
            wins[i0:i0 + B] = (blk > p[:, i0:i0 + B].T).sum(1)
        order = np.argsort(-wins, kind="stable")
    else:
        wins = oracle.wins(p)
    out = {
        "workload": f"synthesized preferences, n={n}, seed={seed}, Binomial(65535, clip(0.5+0.02z,0,1)), 5% forced ties at 32767",
        "n": n,
        "seed": seed,
        "d_sha256": sha(d),
        "p_sha256": sha(p),
        "p_max": int(p.max()),
        "wins_sha256": hashlib.sha256(np.asarray(wins, dtype="<i8").tobytes()).hexdigest(),
        "ranking_sha256": hashlib.sha256(np.asarray(order, dtype="<i8").tobytes()).hexdigest(),
        "ranking_top10": [int(x) for x in order[:10]],
        "oracle": "oracle_closure_tiled T=64",
        "oracle_seconds": round(t2 - t1, 1),
        "oracle_threads": os.cpu_count(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"c3_{n}_seed{seed}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
