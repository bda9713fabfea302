"""Writes the 4096 x 4096 parity fixture: sha256 of the oracle's preference matrix d
(counted ballot by ballot, scaling_elections.py:37-88), of its strongest paths p and of
its ranking, for 4096 seeded impartial-culture ballots over 4096 candidates (seed 2).

Minutes of single-threaded counting; rerun only when the generator or seed changes:
    python tests/golden/make_c2_fixture.py [m] [n] [seed]
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


def sha(a, dt="<u4") -> str:
    return hashlib.sha256(np.ascontiguousarray(a).astype(dt, copy=False).tobytes()).hexdigest()


def main():
    m = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 4096
    seed = int(sys.argv[3]) if len(sys.argv) > 3 else 2
    t0 = time.time()
    R = synth.impartial_culture(m, n, seed)
    d = oracle.build_prefs(list(R))
    t1 = time.time()
    p = oracle.closure_tiled(d, 16, threads=os.cpu_count() or 1)
    t2 = time.time()
    w = oracle.wins(p)
    order = np.argsort(-w, kind="stable")
    out = {
        "workload": f"{m} impartial-culture ballots over {n} candidates, seed {seed}",
        "m": m, "n": n, "seed": seed,
        "rankings_sha256": sha(R, "<i4"),
        "d_sha256": sha(d),
        "p_sha256": sha(p),
        "p_max": int(p.max()),
        "wins_sha256": sha(w, "<i8"),
        "ranking_sha256": sha(order, "<i8"),
        "ranking_top10": [int(x) for x in order[:10]],
        "oracle": "oracle_build_prefs + oracle_closure_tiled T=16",
        "oracle_seconds": {"count": round(t1 - t0, 1), "closure": round(t2 - t1, 1)},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"c2_{m}x{n}_seed{seed}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
