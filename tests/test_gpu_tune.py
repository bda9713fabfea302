"""Launch-shape tuning never changes a result: closures under every table entry (pivot
depth, look-ahead, per width) and batches under every rounds threshold agree with the
oracle and with each other, bit for bit; the autotuner's own sweep runs and verifies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flags(width):
    from scalingelections_amd import _lib

    return {"f16": 0, "u16": _lib.SE_FLAG_NO_F16, "u32": _lib.SE_FLAG_FORCE_U32}[width]


@pytest.mark.parametrize("width", ["f16", "u16", "u32"])
@pytest.mark.parametrize("n", [200, 777])
def test_table_choice_is_used_and_exact(gpu_lib, oracle_mod, width, n):
    """A table installed without flags decides pivot depth and look-ahead; each choice
    gives the oracle's closure."""
    from scalingelections_amd import autotune, strongest_paths_timed

    rng = np.random.default_rng(n + 3)
    d = rng.integers(0, 3000, (n, n)).astype(np.uint32)
    want = oracle_mod.closure_tiled(d, 16)
    try:
        for pivot, la in autotune.candidates(width):
            print(f"tuned closure {width} n={n} pivot={pivot} lookahead={la}", flush=True)
            autotune.install([{"width": width, "max_side": autotune.ANY_SIDE, "pivot": pivot, "lookahead": la}])
            got, t = strongest_paths_timed(d, _flags(width))
            assert t["pivot"] == pivot and t["width_bits"] == (32 if width == "u32" else 16)
            assert t["f16_ordered"] == (1 if width == "f16" else 0)
            assert np.array_equal(got, want), (width, n, pivot, la)
    finally:
        autotune.reset()


def test_table_buckets_select_by_padded_side(gpu_lib, oracle_mod):
    from scalingelections_amd import autotune, strongest_paths_timed

    table = [{"width": "f16", "max_side": 256, "pivot": 32, "lookahead": 1},
             {"width": "f16", "max_side": 512, "pivot": 64, "lookahead": 2}]
    try:
        autotune.install(table)
        # (up to 128 candidates the one-launch small path runs instead of the table)
        for n, pivot in ((129, 32), (256, 32), (257, 64), (512, 64)):
            d = np.random.default_rng(n).integers(0, 900, (n, n)).astype(np.uint32)
            got, t = strongest_paths_timed(d, 0)
            assert t["pivot"] == pivot, (n, t["pivot"])
            assert np.array_equal(got, oracle_mod.closure_tiled(d, 16))
        # beyond every entry of the width: the built-in table decides
        d = np.random.default_rng(5).integers(0, 900, (600, 600)).astype(np.uint32)
        got, t = strongest_paths_timed(d, 0)
        assert t["pivot"] in (32, 64) and np.array_equal(got, oracle_mod.closure_tiled(d, 16))
    finally:
        autotune.reset()


def test_batch_threshold_never_changes_results(gpu_lib, oracle_mod):
    from scalingelections_amd import _lib, autotune, batch

    rng = np.random.default_rng(21)
    C = np.array([8, 64, 65, 128, 129, 192, 193, 256, 257, 300, 384, 385, 512, 513, 700], np.uint32)
    m = rng.integers(50, 400, C.shape[0]).astype(np.uint32)
    phi = rng.uniform(0.5, 1.0, C.shape[0])
    w = batch.make_workload(C, m, phi, seed=21)
    db = batch.DeviceBatch(w)
    ref = None
    try:
        for th in (129, 193, 257, 385, 513, 1025):
            print(f"batch threshold {th}", flush=True)
            _lib.check(_lib.lib().se_tune_batch(th))
            db.run()
            out = db.raw()
            if ref is None:
                ref = out
                paths, ranks = db.results()
                for e in (1, 4, 8, 13):
                    pos = w.election_positions(e)
                    d = oracle_mod.build_prefs([np.argsort(row, kind="stable") for row in pos])
                    p = oracle_mod.closure_tiled(d, 16)
                    assert np.array_equal(paths[e], p), e
                    assert np.array_equal(ranks[e].astype(np.int64), oracle_mod.ranking(p)), e
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), th
    finally:
        autotune.reset()


def test_autotuner_sweep_runs_and_verifies(gpu_lib):
    from scalingelections_amd import autotune

    try:
        entries, meas = autotune.tune_closure(sizes=(256, 384), widths=("f16", "u32"), min_seconds=0.0,
                                              max_reps=2, log=None)
        assert {e["width"] for e in entries} == {"f16", "u32"}
        assert all(e["max_side"] in (256, autotune.ANY_SIDE) for e in entries)
        assert set(meas) == {"f16/256", "f16/384", "u32/256", "u32/384"}
        autotune.install(entries)
        assert autotune.current_table() == entries
        th, times = autotune.tune_batch(elections=64, thresholds=(129, 1025), reps=1, log=None)
        assert th in (129, 1025) and set(times) == {129, 1025}
        assert autotune.batch_rounds_min() == 129
    finally:
        autotune.reset()
