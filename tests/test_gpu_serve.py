"""Serving loop: every request's ranking and strongest-path matrix are bit-identical to
the same election processed alone, under continuous batching, memory-budget admission
and large requests overlapping on a second stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(small, large, n, rate, seed, large_every):
    from scalingelections_amd import serve

    arr = serve.poisson_schedule(rate, n, seed)
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        if large is not None and k % large_every == large_every - 1:
            reqs.append(serve.Request(k, float(arr[k]), large, int(rng.integers(large.elections))))
        else:
            reqs.append(serve.Request(k, float(arr[k]), small, int(rng.integers(small.elections))))
    return reqs


def test_serving_results_equal_solo(gpu_lib):
    from scalingelections_amd import batch, serve

    small = batch.mixed_size_workload(48, seed=21, m_range=(20, 300))
    large = batch.make_workload(np.array([1100, 1536], np.uint32), np.array([60, 90], np.uint32),
                                np.array([0.8, 1.0]), seed=22)
    reqs = _requests(small, large, 240, rate=20000.0, seed=23, large_every=20)
    # budget: room for one large request and a handful of small ones at a time
    budget = max(r.footprint() for r in reqs if r.large) + 8 * max(r.footprint() for r in reqs if not r.large)
    srv = serve.Server(budget_bytes=budget, small_arena_bytes=64 << 20, large_slots=2, large_max_c=1536,
                       large_max_m=90)
    try:
        res = srv.run(reqs)
    finally:
        srv.close()
    summary = res.summary()
    print("serving summary", summary, flush=True)
    assert all(r.done >= r.arrival >= 0 and r.admitted >= 0 for r in res.requests)
    assert res.batches < sum(1 for r in res.requests if not r.large)  # batches held several requests
    solo = {}
    for r in res.requests:
        key = (id(r.pool), r.index)
        if key not in solo:
            solo[key] = serve.solo_result(r.pool, r.index)
        ranking, sha = solo[key]
        assert np.array_equal(r.ranking, ranking), (r.rid, r.C)
        assert r.paths_sha256 == sha, (r.rid, r.C)


def test_admission_rejects_oversize_request(gpu_lib):
    from scalingelections_amd import batch, serve

    small = batch.mixed_size_workload(4, seed=3, m_range=(5, 10))
    r = serve.Request(0, 0.0, small, 0)
    srv = serve.Server(budget_bytes=r.footprint() - 1, small_arena_bytes=1 << 20, large_slots=1, large_max_c=64,
                       large_max_m=8)
    try:
        with pytest.raises(ValueError, match="budget"):
            srv.run([r])
    finally:
        srv.close()
