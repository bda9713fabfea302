"""Pins the CPU oracle before anything is compared against it.

Each oracle function is checked against a source that does not share its code:
golden fixtures (the reference's paradox ballots, the 45-voter textbook example),
brute force by per-threshold reachability, a modified Dijkstra, the reference's own
OpenMP closure compiled from its source (oracle/_ref), and algebraic properties.
A mutation test shows the pins reject plausible mistakes.
"""
import numpy as np
import pytest

from common import (brute_count, brute_widest, closure_is_closed, fill_ballot, load_golden,
                    p0_numpy, python_ranking, schulze45_ballots)


# ---------------------------------------------------------------- golden fixtures
def test_paradox_fixture(oracle_mod):
    g = load_golden("paradox.json")
    for case in g["cases"]:
        ballots = [np.array(b) for b in case["ballots"]]
        d = oracle_mod.build_prefs(ballots)
        assert d.tolist() == case["d"]
        for p in (oracle_mod.closure_serial(d), oracle_mod.closure_tiled(d, 2), oracle_mod.closure_tiled(d, 16)):
            assert p.tolist() == case["p"]
        p = oracle_mod.closure_serial(d)
        assert oracle_mod.wins(p).tolist() == case["wins"]
        assert oracle_mod.ranking(p).tolist() == case["ranking"]
        winner, ranked = oracle_mod.winner_and_ranking(g["candidates"], p)
        assert winner == case["winner"]


def test_schulze45_fixture(oracle_mod):
    ballots, g = schulze45_ballots()
    d = oracle_mod.build_prefs(ballots)
    assert d.tolist() == g["d"]
    p = oracle_mod.closure_serial(d)
    assert p.tolist() == g["p"]
    assert oracle_mod.ranking(p).tolist() == g["ranking"]


def test_fixtures_match_brute_force():
    for name in ("paradox.json",):
        for case in load_golden(name)["cases"]:
            assert brute_widest(np.array(case["d"])).tolist() == case["p"]
    _, g = schulze45_ballots()
    assert brute_widest(np.array(g["d"])).tolist() == g["p"]


# ---------------------------------------------------------------- closure vs independent algorithms
@pytest.mark.parametrize("seed", range(40))
def test_closure_vs_brute_force(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 9))
    hi = int(rng.choice([2, 5, 50, 2**32 - 1]))
    d = rng.integers(0, hi, (n, n), dtype=np.uint64).astype(np.uint32)
    want = brute_widest(d)
    assert np.array_equal(oracle_mod.closure_serial(d), want)
    for T in (1, 2, 3, 4, 16):
        assert np.array_equal(oracle_mod.closure_tiled(d, T), want), T


@pytest.mark.parametrize("n", [16, 50, 129])
def test_closure_vs_dijkstra(oracle_mod, n):
    rng = np.random.default_rng(n)
    d = rng.integers(0, 1000, (n, n)).astype(np.uint32)
    p = oracle_mod.closure_serial(d)
    p0 = oracle_mod.p0(d)
    assert np.array_equal(p0, p0_numpy(d).astype(np.uint32))
    for src in range(0, n, max(1, n // 7)):
        assert np.array_equal(oracle_mod.widest_row(p0, src), p[src])


def test_dijkstra_vs_brute_force(oracle_mod):
    rng = np.random.default_rng(7)
    for _ in range(20):
        n = int(rng.integers(2, 8))
        d = rng.integers(0, 6, (n, n)).astype(np.uint32)
        want = brute_widest(d)
        p0 = p0_numpy(d).astype(np.uint32)
        for s in range(n):
            assert np.array_equal(oracle_mod.widest_row(p0, s), want[s])


@pytest.mark.parametrize("n,T", [(64, 4), (64, 16), (128, 32), (192, 64), (96, 8)])
def test_tiled_vs_reference_openmp(oracle_mod, n, T):
    if not oracle_mod.ref_available():
        pytest.skip("reference OpenMP closure not built (reference tree absent)")
    rng = np.random.default_rng(n * 131 + T)
    d = rng.integers(0, n, (n, n)).astype(np.uint32)
    ref = oracle_mod.ref_closure_openmp(d, T)
    assert np.array_equal(oracle_mod.closure_tiled(d, T), ref)
    assert np.array_equal(oracle_mod.closure_serial(d), ref)


# ---------------------------------------------------------------- properties
@pytest.mark.parametrize("seed", range(6))
def test_closure_properties(oracle_mod, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 60))
    d = rng.integers(0, 30, (n, n)).astype(np.uint32)
    p = oracle_mod.closure_serial(d)
    p0 = p0_numpy(d)
    off = ~np.eye(n, dtype=bool)
    assert (p[off] >= p0[off]).all()
    assert set(np.unique(p).tolist()) <= set(np.unique(p0).tolist()) | {0}
    assert (np.diag(p) == 0).all()
    assert closure_is_closed(p)
    # relabelling equivariance: p(P d P^T) = P p(d) P^T
    perm = rng.permutation(n)
    assert np.array_equal(oracle_mod.closure_serial(d[np.ix_(perm, perm)]), p[np.ix_(perm, perm)])
    # strictly increasing remap commutes with the closure (zero stays zero)
    f = lambda x: (x.astype(np.uint64) * 1000 + 7).astype(np.uint32)  # noqa: E731
    pf = oracle_mod.closure_serial(f(d))
    assert np.array_equal(pf, np.where(p > 0, f(p), 0).astype(np.uint32))
    # tile size and zero padding do not matter
    for T in (3, 7, 16, 64):
        assert np.array_equal(oracle_mod.closure_tiled(d, T), p)


def test_tiled_bounded_rounds_is_prefix(oracle_mod):
    rng = np.random.default_rng(5)
    d = rng.integers(0, 100, (64, 64)).astype(np.uint32)
    full = oracle_mod.closure_tiled(d, 16)
    assert np.array_equal(oracle_mod.closure_tiled(d, 16, max_rounds=4), full)
    part = oracle_mod.closure_tiled(d, 16, max_rounds=1)
    assert (part >= p0_numpy(d).astype(np.uint32)).all()


# ---------------------------------------------------------------- mutation test of the pins
def _serial_mutant(d, kind):
    d = np.asarray(d, dtype=np.int64)
    n = d.shape[0]
    if kind == "ge":
        p = np.where(d >= d.T, d, 0)
    else:
        p = np.where(d > d.T, d, 0)
    np.fill_diagonal(p, 0)
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            for k in range(n):
                if i == k or j == k:
                    continue
                if kind == "maxmin":
                    p[j, k] = min(p[j, k], max(p[j, i], p[i, k]))
                elif kind == "transposed":
                    p[j, k] = max(p[j, k], min(p[i, j], p[i, k]))
                elif kind == "no_max":
                    p[j, k] = min(p[j, i], p[i, k])
                else:
                    p[j, k] = max(p[j, k], min(p[j, i], p[i, k]))
    return p


@pytest.mark.parametrize("kind", ["ge", "maxmin", "transposed", "no_max"])
def test_pins_reject_planted_mistakes(kind):
    ballots, g = schulze45_ballots()
    caught = _serial_mutant(np.array(g["d"]), kind).tolist() != g["p"]
    for case in load_golden("paradox.json")["cases"]:
        caught |= _serial_mutant(np.array(case["d"]), kind).tolist() != case["p"]
    rng = np.random.default_rng(0)
    for _ in range(10):
        d = rng.integers(0, 6, (6, 6))
        caught |= not np.array_equal(_serial_mutant(d, kind), brute_widest(d.astype(np.uint32)))
    assert caught, kind
    # the unmutated restatement passes the same pins
    assert _serial_mutant(np.array(g["d"]), "ok").tolist() == g["p"]


# ---------------------------------------------------------------- counter and builder
def test_build_prefs_vs_brute_force(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(30):
        n = int(rng.integers(1, 9))
        m = int(rng.integers(0, 12))
        ballots = []
        for _ in range(m):
            L = int(rng.integers(1, n + 1))
            r = rng.permutation(n)[:L]
            if rng.random() < 0.2 and L > 1:
                r[-1] = r[0]  # a repeated id, counted the reference's way
            ballots.append(r)
        if not ballots:
            continue
        nn = max(1, max(int(np.max(b)) + 1 for b in ballots))
        want = brute_count([fill_ballot(b, nn) for b in ballots], nn)
        assert np.array_equal(oracle_mod.build_prefs(ballots), want)


def test_counter_identities(oracle_mod):
    rng = np.random.default_rng(11)
    n, m = 23, 57
    R = [rng.permutation(n) for _ in range(m)]
    d = oracle_mod.build_prefs(R).astype(np.int64)
    off = ~np.eye(n, dtype=bool)
    assert ((d + d.T)[off] == m).all() and (np.diag(d) == 0).all()
    pos = np.array([np.argsort(r) for r in R])
    assert np.array_equal(d.sum(1), (n - 1 - pos).sum(0))


def test_populate_single_ballot(oracle_mod):
    p = np.zeros((4, 4), dtype=np.uint32)
    oracle_mod.populate(p, [2, 0, 3, 1])
    want = np.zeros((4, 4), dtype=np.uint32)
    for a, b in [(2, 0), (2, 3), (2, 1), (0, 3), (0, 1), (3, 1)]:
        want[a, b] = 1
    assert np.array_equal(p, want)


# ---------------------------------------------------------------- ranking
def test_ranking_is_python_stable_sort(oracle_mod):
    rng = np.random.default_rng(9)
    for _ in range(20):
        n = int(rng.integers(1, 40))
        d = rng.integers(0, 3, (n, n)).astype(np.uint32)  # many ties
        p = oracle_mod.closure_serial(d)
        w = oracle_mod.wins(p)
        wins_py = [sum(1 for j in range(n) if i != j and p[i, j] > p[j, i]) for i in range(n)]
        assert w.tolist() == wins_py
        assert oracle_mod.ranking(p).tolist() == python_ranking(wins_py)
