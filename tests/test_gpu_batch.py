"""Batched mixed-size elections: every election's strongest paths and ranking from the
batched launches equal the per-launch form's and the oracle's, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE_C = [1, 2, 3, 7, 8, 9, 31, 32, 33, 63, 64, 65, 127, 128, 129, 200, 255, 256, 257, 300, 511, 512, 513, 700]


def _oracle_election(oracle_mod, pos, C):
    m = pos.shape[0]
    if m == 0:
        d = np.zeros((C, C), np.uint32)
    else:
        rankings = [np.argsort(row, kind="stable") for row in pos]
        d = oracle_mod.build_prefs(rankings)
        assert d.shape == (C, C)
    p = oracle_mod.closure_tiled(d, 16)
    return d, p, oracle_mod.ranking(p)


def _check(oracle_mod, w, label, sample=None):
    """Batched (large elections closed by shared per-round launches), batched with every
    election in one workgroup, and one launch sequence per election must agree bit for bit;
    a sample (or all) is also checked against the oracle."""
    from scalingelections_amd import batch

    db = batch.DeviceBatch(w)
    db.run(per_launch=False)
    pb, rb = db.results()
    db.run(per_launch=False, one_group=True)
    pg, rg = db.results()
    db.run(per_launch=True)
    pl, rl = db.results()
    idx = range(w.elections) if sample is None else sample
    for e in range(w.elections):
        assert np.array_equal(pb[e], pl[e]), (label, "paths batched vs per-launch", e, int(w.C[e]))
        assert np.array_equal(rb[e], rl[e]), (label, "ranking batched vs per-launch", e, int(w.C[e]))
        assert np.array_equal(pg[e], pl[e]), (label, "paths one-group vs per-launch", e, int(w.C[e]))
        assert np.array_equal(rg[e], rl[e]), (label, "ranking one-group vs per-launch", e, int(w.C[e]))
    for e in idx:
        _, p, r = _oracle_election(oracle_mod, w.election_positions(e), int(w.C[e]))
        assert np.array_equal(pb[e], p), (label, "paths vs oracle", e, int(w.C[e]))
        assert np.array_equal(rb[e].astype(np.int64), r), (label, "ranking vs oracle", e, int(w.C[e]))


def test_batch_edge_sizes(gpu_lib, oracle_mod):
    from scalingelections_amd import batch

    rng = np.random.default_rng(0)
    C = np.array(EDGE_C, dtype=np.uint32)
    m = rng.integers(1, 120, C.shape[0]).astype(np.uint32)
    phi = rng.uniform(0.5, 1.0, C.shape[0])
    print("batch edge sizes", C.tolist(), flush=True)
    _check(oracle_mod, batch.make_workload(C, m, phi, seed=11), "edge")


def test_batch_zero_and_single_voter(gpu_lib, oracle_mod):
    from scalingelections_amd import batch

    C = np.array([5, 70, 1, 130, 2], dtype=np.uint32)
    m = np.array([0, 1, 0, 1, 3], dtype=np.uint32)
    _check(oracle_mod, batch.make_workload(C, m, np.ones(5), seed=12), "tiny")


def test_batch_many_voters_every_packing(gpu_lib, oracle_mod):
    """Counter packing per election: bytes (C <= 128), 16-bit halves (C <= 1024 with at
    most 65534 voters) and one int32 per voter (more voters than a 16-bit slice holds);
    70000 voters also take the batch's 32-bit closure."""
    from scalingelections_amd import batch

    C = np.array([100, 200, 200, 9], dtype=np.uint32)
    m = np.array([70000, 30000, 70000, 70000], dtype=np.uint32)
    print("batch many voters", C.tolist(), m.tolist(), flush=True)
    _check(oracle_mod, batch.make_workload(C, m, np.array([0.9, 0.95, 0.97, 0.6]), seed=14), "voters")


def test_batch_16bit_integer_closure(gpu_lib, oracle_mod):
    """More than 30720 voters but at most 65535: the batch closure runs 16-bit integer
    tiles instead of the f16-ordered ones (whose offset needs values up to 30720)."""
    from scalingelections_amd import batch

    C = np.array([100, 200, 129, 17], dtype=np.uint32)
    m = np.array([40000, 33000, 31000, 30721], dtype=np.uint32)
    print("batch 16-bit integer", C.tolist(), m.tolist(), flush=True)
    _check(oracle_mod, batch.make_workload(C, m, np.array([0.8, 0.99, 0.7, 0.95]), seed=15), "u16")


def test_batch_mixed_sample(gpu_lib, oracle_mod):
    """200 elections from the mixed-size distribution (smaller voter counts); all compared
    batched vs per-launch, 40 against the oracle."""
    from scalingelections_amd import batch

    w = batch.mixed_size_workload(200, seed=13, m_range=(20, 400))
    _check(oracle_mod, w, "mixed", sample=range(0, 200, 5))


def test_batch_rejects_oversize_election(gpu_lib):
    from scalingelections_amd import batch

    w = batch.make_workload(np.array([1025], np.uint32), np.array([1], np.uint32), np.ones(1), seed=1)
    db = batch.DeviceBatch(w)
    with pytest.raises(RuntimeError, match="1024 candidates"):
        db.run()


def test_batch_int32_and_uint16_positions_agree(gpu_lib, oracle_mod):
    """SE_BATCH_POS16 (uint16 positions, converted on the device) and int32 positions give
    the same results, batched and per launch; the conversion rejects values above 65535."""
    from scalingelections_amd import _lib, batch, devmem

    rng = np.random.default_rng(3)
    C = np.array([5, 64, 100, 128, 129, 200, 300, 512, 700], np.uint32)
    m = rng.integers(1, 3000, C.shape[0]).astype(np.uint32)
    w = batch.make_workload(C, m, rng.uniform(0.5, 1.0, C.shape[0]), seed=3)
    a, b = batch.DeviceBatch(w, pos16=True), batch.DeviceBatch(w, pos16=False)
    for per_launch in (False, True):
        a.run(per_launch=per_launch)
        b.run(per_launch=per_launch)
        ra, rb = a.raw(), b.raw()
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), per_launch
    paths, ranks = a.results()
    for e in (0, 3, 7):
        pos = w.election_positions(e)
        d = oracle_mod.build_prefs([np.argsort(row, kind="stable") for row in pos])
        assert np.array_equal(paths[e], oracle_mod.closure_tiled(d, 16)), e
    bad = np.array([0, 1, 65535, 65536, 7], np.int32)
    src = devmem.DeviceBuffer.from_array(bad)
    dst = devmem.DeviceBuffer(bad.size * 2)
    rc = _lib.lib().se_positions16_device(src.ptr, bad.size, dst.ptr, None)
    assert rc == _lib.SE_ERR_INVALID and "0..65535" in _lib.last_error()
    ok = np.array([0, 1, 65535, 1023], np.int32)
    src2 = devmem.DeviceBuffer.from_array(ok)
    _lib.check(_lib.lib().se_positions16_device(src2.ptr, ok.size, dst.ptr, None))
    assert np.array_equal(dst.download(np.empty(ok.size, np.uint16)), ok.astype(np.uint16))
