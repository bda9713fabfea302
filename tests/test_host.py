"""CPU-only tests of the product: the C ABI library loads and exports every declared
symbol, argument validation and error texts, the explicit CPU branch (allow_gpu=False)
and the serial baseline against the oracle, interruption, input conversion, generators.
No GPU compute is called here."""
import ctypes
import os
import re
import signal
import threading
import time

import numpy as np
import pytest

from common import load_golden, p0_numpy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            syms |= set(re.findall(r"\b(se_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    from scalingelections_amd import _lib

    L = _lib.lib()
    declared = _declared_symbols()
    assert declared, "no declarations found"
    for s in sorted(declared):
        assert hasattr(L, s), s
    assert set(_lib.EXPORTED) <= declared
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for s in declared:
        getattr(raw, s)


def test_library_is_gfx950_build():
    from scalingelections_amd import _lib

    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_shape_and_dtype_errors():
    from scalingelections_amd import compute_strongest_paths

    with pytest.raises(RuntimeError, match="^Number of dimensions must be two$"):
        compute_strongest_paths(np.zeros(4, dtype=np.uint32))
    with pytest.raises(RuntimeError, match="^Preferences matrix must be square$"):
        compute_strongest_paths(np.zeros((3, 4), dtype=np.uint32))
    for dt in (np.int64, np.float32, np.int32, np.uint64):
        with pytest.raises(TypeError):
            compute_strongest_paths(np.zeros((4, 4), dtype=dt))
    with pytest.raises(TypeError):
        compute_strongest_paths([[0, -1], [1, 0]])


def test_gpu_tile_validation_precedes_device_use():
    from scalingelections_amd import compute_strongest_paths

    d = np.zeros((8, 8), dtype=np.uint32)
    for t in (0, 1, 2, 3, 5, 64, 128, 2**40):
        with pytest.raises(RuntimeError, match="^Unsupported tile size$"):
            compute_strongest_paths(d, allow_gpu=True, tile_size=t)


def test_cpu_tile_rules():
    from scalingelections_amd import compute_strongest_paths

    with pytest.raises(RuntimeError, match="^Number of candidates should be at least 4, ideally divisible by 4$"):
        compute_strongest_paths(np.zeros((3, 3), dtype=np.uint32))
    with pytest.raises(RuntimeError, match="^Tile size should be less than or equal to the number of candidates$"):
        compute_strongest_paths(np.zeros((8, 8), dtype=np.uint32), tile_size=16)
    with pytest.raises(RuntimeError, match="^Unsupported tile size$"):
        compute_strongest_paths(np.zeros((8, 8), dtype=np.uint32), tile_size=6)


@pytest.mark.parametrize("n,tile", [(4, 0), (5, 4), (37, 8), (64, 16), (100, 32), (130, 64), (128, 128)])
def test_cpu_branch_matches_oracle(oracle_mod, n, tile):
    from scalingelections_amd import compute_strongest_paths

    rng = np.random.default_rng(n + tile)
    d = rng.integers(0, 3 * n, (n, n)).astype(np.uint32)
    want = oracle_mod.closure_serial(d)
    got = compute_strongest_paths(d, allow_gpu=False, tile_size=tile)
    assert got.dtype == np.uint32 and got.shape == (n, n)
    assert np.array_equal(got, want)


def test_cpu_branch_golden():
    from scalingelections_amd import compute_strongest_paths, compute_strongest_paths_serial

    for case in load_golden("paradox.json")["cases"]:
        d = np.array(case["d"], dtype=np.uint32)
        assert compute_strongest_paths(d).tolist() == case["p"]
        assert compute_strongest_paths_serial(d).tolist() == case["p"]


def test_serial_baseline_matches_oracle(oracle_mod):
    from scalingelections_amd import compute_strongest_paths_serial

    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 64, 150):
        d = rng.integers(0, 2**32 - 1, (n, n), dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(compute_strongest_paths_serial(d), oracle_mod.closure_serial(d))


def test_input_conversion_and_no_mutation(oracle_mod):
    from scalingelections_amd import compute_strongest_paths

    rng = np.random.default_rng(2)
    d16 = rng.integers(0, 50, (20, 20)).astype(np.uint16)
    want = oracle_mod.closure_serial(d16.astype(np.uint32))
    assert np.array_equal(compute_strongest_paths(d16, tile_size=4), want)  # safe cast u16 -> u32
    big = rng.integers(0, 50, (40, 40)).astype(np.uint32)
    view = big[::2, ::2]  # non-contiguous: copied
    before = big.copy()
    assert np.array_equal(compute_strongest_paths(view, tile_size=4), oracle_mod.closure_serial(np.ascontiguousarray(view)))
    assert np.array_equal(big, before)
    assert np.array_equal(compute_strongest_paths(d16.astype(np.uint32).tolist(), tile_size=4), want)


def test_cpu_branch_stops_on_sigint():
    """SIGINT during the CPU branch -> RuntimeError("Stopped by signal"); the previous
    handler is restored afterwards."""
    from scalingelections_amd import compute_strongest_paths

    d = np.random.default_rng(0).integers(0, 1000, (2048, 2048)).astype(np.uint32)
    prev = signal.getsignal(signal.SIGINT)

    def fire():
        time.sleep(0.05)
        os.kill(os.getpid(), signal.SIGINT)

    th = threading.Thread(target=fire)
    th.start()
    t0 = time.time()
    with pytest.raises(RuntimeError, match="^Stopped by signal$"):
        compute_strongest_paths(d, tile_size=4)  # 8.6e9 cell updates: seconds, far above 50 ms
    th.join()
    assert time.time() - t0 < 5.0
    assert signal.getsignal(signal.SIGINT) is prev


def test_reduce_sequential_float():
    from scalingelections_amd import reduce

    x = np.random.default_rng(4).standard_normal(1001).astype(np.float32)
    acc = np.float32(0)
    for v in x:
        acc = np.float32(acc + v)
    assert reduce(x) == float(acc)
    assert reduce(np.zeros(0, dtype=np.float32)) == 0.0


def test_builder_positions_host_logic(oracle_mod):
    """Ballot -> position conversion reproduces the reference's fill rule (counted here by
    the oracle from the positions' implied orders, no GPU)."""
    from scalingelections_amd import builder

    rng = np.random.default_rng(8)
    for _ in range(20):
        n = int(rng.integers(1, 12))
        ballots = [rng.permutation(n)[: int(rng.integers(1, n + 1))] for _ in range(int(rng.integers(1, 9)))]
        nn = builder.count_candidates(ballots)
        pos = builder.rankings_to_positions(ballots, nn)
        assert pos is not None
        orders = [np.argsort(row, kind="stable") for row in pos]
        assert np.array_equal(oracle_mod.build_prefs(orders), oracle_mod.build_prefs(ballots))
    assert builder.rankings_to_positions([np.array([0, 1, 0])], 2) is None


def test_synthesized_preferences_shape_and_rules():
    from scalingelections_amd import synth

    d = synth.synthesized_preferences(300, seed=3)
    assert d.dtype == np.uint32
    off = ~np.eye(300, dtype=bool)
    ties = (d == 32767) & (d.T == 32767)
    assert (((d.astype(np.int64) + d.T) == 65535) | ties)[off].all()
    assert 0.03 < ties[off].mean() < 0.07
    assert np.array_equal(d, synth.synthesized_preferences(300, seed=3, workers=1))
    assert p0_numpy(d).max() <= 65535


def test_impartial_culture_is_permutations():
    from scalingelections_amd import synth

    R = synth.impartial_culture(50, 30, seed=1)
    assert (np.sort(R, axis=1) == np.arange(30)).all()
    pos = synth.positions_from_rankings(R)
    assert (np.take_along_axis(pos, R.astype(np.intp), axis=1) == np.arange(30)).all()


def test_cli_flags_and_defaults():
    from scalingelections_amd import cli

    a = cli.parse_args([])
    assert (a.num_voters, a.num_candidates, a.run_cpu, a.run_gpu, a.no_serial) == (0, 256, False, False, False)
    assert (a.cpu_tile_size, a.gpu_tile_size, a.warmup, a.repeat) == (16, 32, 1, 1)
    assert cli.format_throughput(3.5e12) == "3.5 Tcells³/s"
    assert cli.format_time(0.5) == "500 ms" and cli.format_time(2.0) == "2.00 s"


def test_drop_in_module_names():
    import scaling_elections

    for name in ("compute_strongest_paths", "log_gpus", "reduce", "build_pairwise_preferences",
                 "get_winner_and_ranking"):
        assert callable(getattr(scaling_elections, name))


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "scalingelections_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text and "libref" not in text, f


def test_mallows_generator_is_deterministic_and_permutations():
    from scalingelections_amd import batch

    C = np.array([1, 5, 64, 65, 130], np.uint32)
    m = np.array([3, 50, 20, 7, 4], np.uint32)
    phi = np.array([1.0, 0.5, 0.9, 0.7, 1.0])
    ref = np.concatenate([np.random.default_rng(i).permutation(int(c)) for i, c in enumerate(C)]).astype(np.int32)
    a = batch.mallows_positions(C, m, phi, ref, seed=9, threads=1)
    b = batch.mallows_positions(C, m, phi, ref, seed=9, threads=7)
    assert np.array_equal(a, b)
    o = 0
    for c, mm in zip(C, m):
        blk = a[o:o + int(c) * int(mm)].reshape(int(mm), int(c))
        assert (np.sort(blk, axis=1) == np.arange(int(c))).all()
        o += int(c) * int(mm)


@pytest.mark.parametrize("phi", [0.5, 1.0])
def test_mallows_generator_matches_closed_form(phi):
    """Mallows with dispersion phi: P(ballot = reference order) = 1 / prod_k Z_k with
    Z_k = sum_{t<=k} phi^t, and the ballot's inversion count has mean sum_k E[off_k]."""
    from scalingelections_amd import batch

    C, M = 5, 40000
    ref = np.array([3, 0, 4, 1, 2], np.int32)
    pos = batch.mallows_positions(np.array([C], np.uint32), np.array([M], np.uint32), np.array([phi]), ref, seed=5)
    pos = pos.reshape(M, C)
    rank_of_ref = pos[:, ref]  # place of the k-th reference item
    Z = [sum(phi ** t for t in range(k + 1)) for k in range(C)]
    p_id = 1.0 / np.prod(Z)
    got = float((rank_of_ref == np.arange(C)).all(axis=1).mean())
    assert abs(got - p_id) < 5 * np.sqrt(p_id * (1 - p_id) / M)
    inv = sum((rank_of_ref[:, i] > rank_of_ref[:, j]).astype(np.int64) for i in range(C) for j in range(i + 1, C))
    mean_off = sum(sum(t * phi ** t for t in range(k + 1)) / Z[k] for k in range(C))
    assert abs(inv.mean() - mean_off) < 0.05


def _no_device():
    from scalingelections_amd import api

    return api.device_count() == 0


def test_gpu_paths_fail_loudly_without_a_device():
    """No silent CPU fallback: every GPU entry point raises when no device is visible."""
    from scalingelections_amd import build_pairwise_preferences, compute_strongest_paths, get_winner_and_ranking

    if not _no_device():
        pytest.skip("a HIP device is visible")
    d = np.zeros((8, 8), np.uint32)
    with pytest.raises(RuntimeError, match="No HIP devices available"):
        compute_strongest_paths(d, allow_gpu=True, tile_size=32)
    with pytest.raises(RuntimeError, match="No HIP devices available"):
        build_pairwise_preferences([np.array([0, 1, 2])])
    with pytest.raises(RuntimeError, match="No HIP devices available"):
        get_winner_and_ranking([0, 1], np.zeros((2, 2), np.uint32))


def test_ballot_validation_before_device_use():
    from scalingelections_amd import _lib

    ids = np.array([0, 5], np.int32)
    off = np.array([0, 2], np.int64)
    out = np.zeros((3, 3), np.uint32)
    rc = _lib.lib().se_count_rankings(ids.ctypes.data, off.ctypes.data, 1, 3, out.ctypes.data, 3)
    assert rc == _lib.SE_ERR_INVALID and "[0, n)" in _lib.last_error()
    bad_off = np.array([2, 0], np.int64)
    rc = _lib.lib().se_count_rankings(ids.ctypes.data, bad_off.ctypes.data, 1, 6, out.ctypes.data, 3)
    assert rc == _lib.SE_ERR_INVALID
    with pytest.raises(ValueError):
        from scalingelections_amd import builder

        builder.rankings_to_positions([np.array([0, -1])], 2)


def test_tuning_table_api():
    """The launch-shape table: built-in entries cover every width and side, installs and
    resets round-trip, invalid entries are rejected with a message, and the bucket rule
    (the smallest measured side >= N decides) merges equal neighbours."""
    from scalingelections_amd import _lib, autotune

    autotune.reset()
    base = autotune.current_table()
    for w in ("f16", "u16", "u32"):
        assert any(e["width"] == w and e["max_side"] == autotune.ANY_SIDE for e in base), w
        assert all(e["pivot"] in (32, 64) and e["lookahead"] in (0, 1, 2) for e in base)
    assert autotune.candidates("u32") == ((32, 0), (32, 1), (64, 0), (64, 1))
    assert autotune.batch_rounds_min() == 129
    try:
        mine = [{"width": "u32", "max_side": 512, "pivot": 64, "lookahead": 1}]
        autotune.install(mine, 1025)
        assert autotune.current_table() == mine and autotune.batch_rounds_min() == 1025
        L = _lib.lib()
        bad = (_lib.se_tune_entry * 1)(_lib.se_tune_entry(0, 512, 48, 0))
        assert L.se_tune_closure(bad, 1) == _lib.SE_ERR_INVALID and "pivot 32 or 64" in _lib.last_error()
        bad[0] = _lib.se_tune_entry(3, 512, 32, 0)
        assert L.se_tune_closure(bad, 1) == _lib.SE_ERR_INVALID
        bad[0] = _lib.se_tune_entry(0, 0, 32, 0)
        assert L.se_tune_closure(bad, 1) == _lib.SE_ERR_INVALID
        for w, piv, la in ((2, 64, 2), (0, 32, 2), (1, 64, 3)):  # one launch per round: 16-bit, pivot 64
            bad[0] = _lib.se_tune_entry(w, 512, piv, la)
            assert L.se_tune_closure(bad, 1) == _lib.SE_ERR_INVALID, (w, piv, la)
        assert L.se_tune_closure(None, 65) == _lib.SE_ERR_INVALID
        for th in (1, 128, 1026):
            assert L.se_tune_batch(th) == _lib.SE_ERR_INVALID and "129..1025" in _lib.last_error()
        assert autotune.current_table() == mine and autotune.batch_rounds_min() == 1025
    finally:
        autotune.reset()
    assert autotune.current_table() == base and autotune.batch_rounds_min() == 129
    got = autotune.buckets({1024: (64, 1), 256: (32, 1), 512: (32, 1), 2048: (32, 0)}, "u16")
    assert got == [{"width": "u16", "max_side": 512, "pivot": 32, "lookahead": 1},
                   {"width": "u16", "max_side": 1024, "pivot": 64, "lookahead": 1},
                   {"width": "u16", "max_side": autotune.ANY_SIDE, "pivot": 32, "lookahead": 0}]
    assert autotune.padded_side(1) == 128 and autotune.padded_side(256) == 256 and autotune.padded_side(257) == 384


def test_tuning_table_save_load_roundtrip(tmp_path, monkeypatch):
    from scalingelections_amd import api, autotune

    monkeypatch.setattr(api, "device_info", lambda device=0: {"name": "test"})
    entries = [{"width": "f16", "max_side": 1024, "pivot": 32, "lookahead": 0},
               {"width": "f16", "max_side": autotune.ANY_SIDE, "pivot": 64, "lookahead": 1}]
    path = str(tmp_path / "tune.json")
    try:
        autotune.save(path, entries, 257, {"f16/256": {"T32": 0.1}})
        autotune.reset()
        doc = autotune.load(path)
        assert doc["device"] == {"name": "test"}
        assert autotune.current_table() == entries and autotune.batch_rounds_min() == 257
    finally:
        autotune.reset()
