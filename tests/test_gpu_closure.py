"""GPU parity: the HIP closure, counter and ranking against the CPU oracle, bit-exact.

Sizes here finish on the oracle in seconds; the 32768-candidate workload is covered by
test_gpu_large.py (hash fixture + independent per-row Dijkstra checks)."""
import ctypes

import numpy as np
import pytest

from common import load_golden, schulze45_ballots

pytestmark = pytest.mark.gpu


def _gpu(d, tile=32, tma=False):
    from scalingelections_amd import compute_strongest_paths

    return compute_strongest_paths(d, allow_gpu=True, tile_size=tile, allow_tma=tma)


def test_golden_fixtures_gpu(gpu_lib):
    for case in load_golden("paradox.json")["cases"]:
        d = np.array(case["d"], dtype=np.uint32)
        for t in (4, 8, 16, 32):
            assert _gpu(d, t).tolist() == case["p"]
    _, g = schulze45_ballots()
    assert _gpu(np.array(g["d"], dtype=np.uint32)).tolist() == g["p"]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 31, 32, 33, 100, 127, 128, 129, 255, 256, 257, 300, 513, 777, 1000])
def test_closure_random_sizes(gpu_lib, oracle_mod, n):
    print(f"closure parity n={n}", flush=True)
    rng = np.random.default_rng(1000 + n)
    d = rng.integers(0, max(2, 2 * n), (n, n)).astype(np.uint32)
    want = oracle_mod.closure_tiled(d, 16)
    got = _gpu(d)
    assert got.dtype == np.uint32 and got.shape == (n, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tile", [4, 8, 16, 32])
def test_tile_size_and_tma_do_not_change_result(gpu_lib, oracle_mod, tile):
    rng = np.random.default_rng(tile)
    d = rng.integers(0, 500, (200, 200)).astype(np.uint32)
    want = oracle_mod.closure_serial(d)
    assert np.array_equal(_gpu(d, tile), want)
    assert np.array_equal(_gpu(d, tile, tma=True), want)


def _timed(d, flags):
    from scalingelections_amd import strongest_paths_timed

    return strongest_paths_timed(d, flags)


@pytest.mark.parametrize("hi,width", [(1000, 16), (65536, 16), (65537, 32), (2**32 - 1, 32)])
def test_width_selection_is_exact(gpu_lib, oracle_mod, hi, width):
    """u16 when max(p0) <= 65535 (boundary included), u32 above; forced widths agree."""
    from scalingelections_amd import _lib

    rng = np.random.default_rng(hi % 1000)
    n = 300
    d = rng.integers(0, hi, (n, n), dtype=np.uint64).astype(np.uint32)
    if hi == 65536:
        d[3, 5], d[5, 3] = 65535, 0  # max(p0) hits the boundary exactly
    if hi == 65537:
        d[3, 5], d[5, 3] = 65536, 0
    want = oracle_mod.closure_tiled(d, 16)
    got, t = _timed(d, 0)
    assert t["width_bits"] == width
    assert np.array_equal(got, want)
    got32, t32 = _timed(d, _lib.SE_FLAG_FORCE_U32)
    assert t32["width_bits"] == 32 and np.array_equal(got32, want)
    if width == 16:
        got16, t16 = _timed(d, _lib.SE_FLAG_FORCE_U16)
        assert t16["width_bits"] == 16 and np.array_equal(got16, want)
    else:
        with pytest.raises(RuntimeError, match="exceed 65535"):
            _timed(d, _lib.SE_FLAG_FORCE_U16)


@pytest.mark.parametrize("schedule", ["fused", "lookahead", "serial"])
@pytest.mark.parametrize("width", ["f16", "u16", "u32"])
@pytest.mark.parametrize("pivot", [32, 64])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 511, 640, 1500])
def test_every_width_and_pivot_depth(gpu_lib, oracle_mod, schedule, width, pivot, n):
    from scalingelections_amd import _lib

    rng = np.random.default_rng(n * 7 + pivot)
    hi = 20000 if width == "f16" else 60000
    d = rng.integers(0, hi, (n, n)).astype(np.uint32)
    flags = {"f16": 0, "u16": _lib.SE_FLAG_FORCE_U16 | _lib.SE_FLAG_NO_F16, "u32": _lib.SE_FLAG_FORCE_U32}[width]
    flags |= _lib.SE_FLAG_PIVOT32 if pivot == 32 else _lib.SE_FLAG_PIVOT64
    flags |= {"serial": _lib.SE_FLAG_NO_LOOKAHEAD, "lookahead": _lib.SE_FLAG_LOOKAHEAD,
              "fused": _lib.SE_FLAG_LOOKAHEAD_FUSED}[schedule]
    got, t = _timed(d, flags)
    assert (t["pivot"], t["width_bits"]) == (pivot, 32 if width == "u32" else 16)
    assert t["f16_ordered"] == (1 if (width == "f16" and n > 1) else 0)
    assert t["rounds"] == (n + pivot - 1) // pivot
    assert np.array_equal(got, oracle_mod.closure_tiled(d, 16))


def _ranged(n, lo, hi, seed):
    """d whose nonzero p0 values are exactly the upper triangle, spanning [lo, hi]."""
    rng = np.random.default_rng(seed)
    d = np.triu(rng.integers(lo, hi + 1, (n, n), dtype=np.uint64), 1).astype(np.uint32)
    d[0, 1], d[0, 2] = lo, hi
    return d


@pytest.mark.parametrize("lo,hi,f16", [
    (1, 30721, True),            # span 30720: top value lands on 0x7C00 (+inf), still ordered
    (1, 30722, False),           # span 30721: one past the f16-ordered range -> integer min/max
    (34815, 65535, True),        # span 30720 at the top of the 16-bit range
    (40000, 40100, True),
    (7, 7, True),                # a single nonzero value
])
def test_f16_ordered_mode_boundaries(gpu_lib, oracle_mod, lo, hi, f16):
    from scalingelections_amd import _lib

    for n in (130, 300):
        d = _ranged(n, lo, hi, lo + n)
        want = oracle_mod.closure_tiled(d, 16)
        got, t = _timed(d, 0)
        assert t["width_bits"] == 16 and t["f16_ordered"] == (1 if f16 else 0)
        assert np.array_equal(got, want)
        got2, t2 = _timed(d, _lib.SE_FLAG_NO_F16)
        assert t2["f16_ordered"] == 0 and np.array_equal(got2, want)
        got3, t3 = _timed(d, _lib.SE_FLAG_PIVOT32)
        assert t3["pivot"] == 32 and np.array_equal(got3, want)


@pytest.mark.parametrize("hi", [300, 70000, 2**32 - 1])
def test_small_elections_one_launch(gpu_lib, oracle_mod, hi):
    """Up to 128 candidates the closure is one workgroup and one launch (32-bit in
    registers, no width decision); values below and above 65535 (16- and 32-bit transit of
    the host entry point), against the oracle; a shape flag selects the blocked path."""
    from scalingelections_amd import _lib

    for n in (1, 2, 3, 16, 17, 64, 100, 127, 128):
        rng = np.random.default_rng(n * 31 + hi % 97)
        d = rng.integers(0, hi, (n, n), dtype=np.uint64).astype(np.uint32)
        want = oracle_mod.closure_serial(d)
        got, t = _timed(d, 0)
        assert (t["pivot"], t["rounds"], t["width_bits"]) == (1, n, 32), (n, t)
        assert np.array_equal(got, want), (n, hi)
        got2, t2 = _timed(d, _lib.SE_FLAG_FORCE_U32)
        assert t2["pivot"] in (32, 64) and np.array_equal(got2, want), (n, hi)
    n = 129  # first size on the blocked path
    d = np.random.default_rng(5).integers(0, hi, (n, n), dtype=np.uint64).astype(np.uint32)
    got, t = _timed(d, 0)
    assert t["pivot"] in (32, 64) and np.array_equal(got, oracle_mod.closure_serial(d))


def test_degenerate_inputs(gpu_lib, oracle_mod):
    for d in (np.zeros((64, 64), np.uint32), np.full((70, 70), 7, np.uint32),
              np.triu(np.full((130, 130), 9, np.uint32), 1), np.eye(40, dtype=np.uint32) * 5):
        assert np.array_equal(_gpu(d), oracle_mod.closure_serial(d))
    assert _gpu(np.zeros((0, 0), np.uint32)).shape == (0, 0)


def test_repeat_calls_reuse_scratch(gpu_lib, oracle_mod):
    rng = np.random.default_rng(77)
    for n in (600, 100, 600, 1200, 64):
        d = rng.integers(0, 90000, (n, n)).astype(np.uint32)
        assert np.array_equal(_gpu(d), oracle_mod.closure_tiled(d, 16))


def test_device_entry_point_with_caller_buffers(gpu_lib, oracle_mod):
    """se_closure_device on caller-owned device memory, padded output stride, caller stream."""
    from scalingelections_amd import _lib, devmem

    rng = np.random.default_rng(5)
    n = 333
    d = rng.integers(0, 4000, (n, n)).astype(np.uint32)
    want = oracle_mod.closure_serial(d)
    dev = devmem.DeviceBuffer.from_array(d)
    out = devmem.DeviceBuffer(n * (n + 3) * 4)
    s = devmem.Stream()
    t = _lib.se_timing()
    rc = _lib.lib().se_closure_device(dev.ptr, n, n, out.ptr, n + 3, _lib.SE_FLAG_TIME_KERNELS,
                                      ctypes.c_void_p(s.handle), ctypes.byref(t))
    _lib.check(rc)
    s.synchronize()
    host = out.download(np.empty((n, n + 3), dtype=np.uint32))
    assert np.array_equal(host[:, :n], want)
    assert t.total_ms > 0 and t.indep_ms > 0 and t.rounds == (n + t.pivot - 1) // t.pivot
    # in place: output over the input buffer
    rc = _lib.lib().se_closure_device(dev.ptr, n, n, dev.ptr, n, 0, None, None)
    _lib.check(rc)
    devmem.device_synchronize()
    assert np.array_equal(dev.download(np.empty((n, n), dtype=np.uint32)), want)


# ---------------------------------------------------------------- counter
def test_counter_c1_pipeline(gpu_lib, oracle_mod):
    """128 candidates x 2000 voters (seed 1): counter, closure, ranking all exact."""
    from scalingelections_amd import build_pairwise_preferences, get_winner_and_ranking, synth

    R = synth.impartial_culture(2000, 128, seed=1)
    ballots = list(R)
    d = build_pairwise_preferences(ballots)
    assert np.array_equal(d, oracle_mod.build_prefs(ballots))
    p = _gpu(d)
    want_p = oracle_mod.closure_serial(d)
    assert np.array_equal(p, want_p)
    cands = list(range(128))
    winner, ranked = get_winner_and_ranking(cands, p)
    w2, r2 = oracle_mod.winner_and_ranking(cands, want_p)
    assert (winner, ranked) == (w2, r2)


@pytest.mark.parametrize("seed", range(6))
def test_counter_incomplete_and_duplicate_ballots(gpu_lib, oracle_mod, seed):
    from scalingelections_amd import build_pairwise_preferences

    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    m = int(rng.integers(1, 200))
    ballots = [rng.permutation(n)[: int(rng.integers(1, n + 1))] for _ in range(m)]
    if seed % 2 == 1:
        ballots[0] = np.array([0, 1, 0, 1])  # repeated ids take the exact scatter path
    d = build_pairwise_preferences(ballots)
    assert np.array_equal(d, oracle_mod.build_prefs(ballots))


@pytest.mark.parametrize("m,n", [(1, 1), (0, 5), (3, 129), (257, 64), (1000, 300), (33, 1000),
                                 (5, 127), (7, 128), (2001, 128), (2001, 129), (130, 256), (3, 4000)])
def test_counter_shapes(gpu_lib, oracle_mod, m, n):
    from scalingelections_amd import count_positions, synth

    R = synth.impartial_culture(m, n, seed=m + n) if m else np.zeros((0, n), np.int32)
    pos = synth.positions_from_rankings(R) if m else np.zeros((0, n), np.int32)
    d = count_positions(pos)
    assert np.array_equal(d, oracle_mod.build_prefs(list(R)) if m else np.zeros((n, n), np.uint32))


def _count_np(pos: np.ndarray) -> np.ndarray:
    """d[i][j] = #{v : pos[v][i] < pos[v][j]} by direct comparison, in voter chunks."""
    m, n = pos.shape
    d = np.zeros((n, n), np.int64)
    for v0 in range(0, m, 512):
        P = pos[v0:v0 + 512]
        d += (P[:, :, None] < P[:, None, :]).sum(0)
    return d.astype(np.uint32)


@pytest.mark.parametrize("budget_planes", [1, 2])
@pytest.mark.parametrize("n", [130, 132])
def test_counter_passes_beyond_16bit_slices(gpu_lib, monkeypatch, budget_planes, n):
    """More voters than one 16-bit-packed slice holds (65534): the counter runs several
    passes that add into d.  A partial-plane budget of 1 or 2 planes forces passes of one
    slice (direct accumulation) and of two slices (partial planes + accumulating reduce,
    one column per load for n = 130, four for n = 132)."""
    from scalingelections_amd import count_positions

    m = 140000
    rng = np.random.default_rng(11)
    pos = np.argsort(rng.random((m, n)), axis=1).astype(np.int32)
    monkeypatch.setenv("SE_COUNT_PARTIAL_BYTES", str(budget_planes * n * n * 4))
    d = count_positions(pos)
    assert np.array_equal(d, _count_np(pos))
    assert np.array_equal(d.astype(np.int64) + d.T, m * (1 - np.eye(n, dtype=np.int64)))


@pytest.mark.parametrize("stride", [0, 4, 3])
def test_count_device_caller_stride(gpu_lib, stride):
    """se_count_device into a caller buffer whose row stride is n + stride: the slice planes
    are summed four columns per load when the stride keeps rows 16-byte aligned (0, 4) and
    one column per load otherwise (3); cells past n in each row are left untouched."""
    from scalingelections_amd import _lib, devmem, synth

    n, m = 256, 3000
    print(f"count device stride n+{stride}", flush=True)
    pos = synth.positions_from_rankings(synth.impartial_culture(m, n, 4))
    ld = n + stride
    pd = devmem.DeviceBuffer.from_array(np.ascontiguousarray(pos, np.int32))
    fill = np.full((n, ld), 0xDEADBEEF, np.uint32)
    dd = devmem.DeviceBuffer.from_array(fill)
    _lib.check(_lib.lib().se_count_device(pd.ptr, m, n, dd.ptr, ld, None, None))
    devmem.device_synchronize()
    got = dd.download(np.empty((n, ld), np.uint32))
    assert np.array_equal(got[:, :n], _count_np(pos))
    assert (got[:, n:] == 0xDEADBEEF).all()


@pytest.mark.parametrize("n,m", [(130, 140000), (100, 70000), (300, 66000), (300, 65535)])
def test_counter_slices_beyond_16bit_range(gpu_lib, n, m):
    """More voters than one packed slice holds (65534 at 16 bits): the voter slices' partial
    planes are summed; also the 8-bit packing (n <= 128) with many voters, counts above
    65535 (32-bit transit back to the host) and at exactly 65535 (16-bit transit)."""
    from scalingelections_amd import count_positions

    print(f"counter slices n={n} m={m}", flush=True)
    rng = np.random.default_rng(11 + n)
    pos = np.argsort(rng.random((m, n)), axis=1).astype(np.int32)
    d = count_positions(pos)
    assert np.array_equal(d, _count_np(pos))
    assert np.array_equal(d.astype(np.int64) + d.T, m * (1 - np.eye(n, dtype=np.int64)))


@pytest.mark.parametrize("n,bad", [(100, 128), (100, -1), (200, 40000), (200, -5)])
def test_counter_rejects_positions_outside_the_packed_range(gpu_lib, n, bad):
    """Positions must be a permutation of 0..n-1; values the packed counter cannot hold
    (>= 128 for n <= 128, >= 32768 for n <= 32768, negative) are reported, not miscounted."""
    from scalingelections_amd import _lib, count_positions

    pos = np.tile(np.arange(n, dtype=np.int32), (9, 1))
    pos[4, 3] = bad
    with pytest.raises(_lib.SEError, match="permutation"):
        count_positions(pos)
    pos[4, 3] = 3
    assert np.array_equal(count_positions(pos), _count_np(pos))


# ---------------------------------------------------------------- ranking
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 1000])
def test_ranking_matches_oracle(gpu_lib, oracle_mod, n):
    from scalingelections_amd import get_winner_and_ranking, wins_and_order

    rng = np.random.default_rng(n)
    d = rng.integers(0, 3, (n, n)).astype(np.uint32)  # heavy ties
    p = oracle_mod.closure_serial(d)
    wins, order = wins_and_order(p)
    assert wins.tolist() == oracle_mod.wins(p).tolist()
    assert order.tolist() == oracle_mod.ranking(p).tolist()
    cands = [f"c{i}" for i in range(n)]
    assert get_winner_and_ranking(cands, p) == oracle_mod.winner_and_ranking(cands, p)


@pytest.mark.parametrize("n", [4000, 4160])
def test_wins_with_multi_block_column_segments(gpu_lib, oracle_mod, n):
    """Sizes where each wins workgroup covers several column blocks, the last one ragged
    (4000) or whole (4160): counts and ranking equal the oracle's on a matrix full of ties."""
    from scalingelections_amd import wins_and_order

    print(f"wins segments n={n}", flush=True)
    p = np.random.default_rng(n).integers(0, 4, (n, n)).astype(np.uint32)
    wins, order = wins_and_order(p)
    assert wins.tolist() == oracle_mod.wins(p).tolist()
    assert order.tolist() == oracle_mod.ranking(p).tolist()


def test_ranking_empty_candidates_raises_like_reference(gpu_lib):
    from scalingelections_amd import get_winner_and_ranking

    with pytest.raises(IndexError):
        get_winner_and_ranking([], np.zeros((0, 0), np.uint32))


def test_log_gpus_reports_gfx950(gpu_lib, capsys):
    from scalingelections_amd import device_info, log_gpus

    log_gpus()
    out = capsys.readouterr().out
    assert "Device 0" in out
    assert device_info(0)["arch"].startswith("gfx950")


@pytest.mark.parametrize("argv,n", [
    (["--num-candidates", "128", "--num-voters", "2000", "--seed", "1"], 128),  # the repo default
    (["--num-candidates", "300", "--num-voters", "0", "--seed", "7"], 300),      # random matrix
])
def test_cli_end_to_end(gpu_lib, oracle_mod, capsys, argv, n):
    """The reference's benchmark CLI (scaling_elections.py:361-565) on the GPU: the serial
    baseline, the tiled CPU branch and both GPU entries run, every one is validated against
    the baseline, and the reported winner and top 5 are the oracle's."""
    from scalingelections_amd import cli

    rc = cli.main(argv + ["--run-cpu", "--run-gpu", "--warmup", "1", "--repeat", "2"])
    out = capsys.readouterr().out
    assert rc == 0, out
    assert out.count("✓ Results validated") == 3 and "✗" not in out, out
    for line in ("→ Serial (C++)", "→ Tiled CPU (C++ threads)", "→ Tiled GPU (HIP)", "(avg of 2)"):
        assert line in out, line
    seed = int(argv[argv.index("--seed") + 1])
    rng = np.random.default_rng(seed)
    if "--num-voters" in argv and argv[argv.index("--num-voters") + 1] == "0":
        d = rng.integers(0, n, (n, n)).astype(np.uint32)
    else:
        m = int(argv[argv.index("--num-voters") + 1])
        d = oracle_mod.build_prefs([rng.permutation(n) for _ in range(m)])
    r = oracle_mod.ranking(oracle_mod.closure_serial(d))
    assert f"Winner: Candidate #{int(r[0])}" in out
    assert f"Top 5:  #{int(r[0])}, #{int(r[1])}, #{int(r[2])}, #{int(r[3])}, #{int(r[4])}" in out
