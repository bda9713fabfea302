"""Benchmark: strongest-path closure throughput on one MI355X, plus the batch and serving
figures of the same metric line.

Headline workload (one step): the full device pipeline of one 32768-candidate election --
strongest-path init (p0) + blocked max-min Floyd-Warshall + output -- on the synthesized
preference matrix (for i < j, d[i][j] ~ Binomial(65535, clip(0.5 + 0.02 z, 0, 1)),
d[j][i] = 65535 - d[i][j], 5% forced ties; seed 3).  The input is resident in HBM before
timing starts.  value = n^3 / step time (TCells/s, the reference's cells^3/s convention).

Also reported, all measured in this run:
  * roofline of the dominant kernel (independent tiles): algorithmic integer ops per launch
    (2 per cell update) / its average launch duration from HIP events recorded around every
    launch on the library stream during the timed steps, against the packed 16-bit VALU peak;
  * end_to_end: the reference's convention (host upload + closure + download) with a
    sha256 check of p against the committed oracle fixture;
  * closure_u32: the same 32768 matrix through the 32-bit closure (SE_FLAG_FORCE_U32, the
    4 GB variant), against the 32-bit VALU bound, with the same sha256 check;
  * closure_by_size: device TCells/s at 4096, 8192 and 16384 candidates;
  * counter: 4096 ballots x 4096 candidates (seed 2), pair-votes/s against the VALU issue
    bound of the packed count (2 instructions per 2 pair-votes), and the whole 4096 x 4096
    election (count, closure, ranking);
  * batch: 4096 mixed-size Mallows elections (C log-uniform in [8, 512], m uniform in
    [100, 10000], seed 4), elections/s batched vs one launch sequence per election, with
    every election's paths and ranking compared between the two;
  * serving: Poisson arrivals from that mix plus 1% 8192 x 8192 Mallows elections, offered
    load set to 50% of the capacity measured above, p50/p99 latency, every result checked
    against the same election processed alone;
  * cpu_baseline: the CPU oracle (tiled restatement of the reference's Numba kernel,
    T = 16) on the 32768 matrix for a bounded number of pivot rounds.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 32768]
                       [--no-cpu] [--no-e2e] [--no-extra]
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# MI355X VALU issue rate: 256 CUs x 64 lanes x 2.4 GHz = 3.93e13 lane-instructions/s.
# Cell updates (c = max(c, min(a, b))) per lane-instruction of the sequence each path uses:
#   f16-ordered 16-bit: 2 v_pk_minimum3_f16 + 1 v_pk_maximum3_f16 per 2 cells x 2 steps -> 4/3
#   integer 16-bit:     1 v_pk_min_u16 + 1 v_pk_max_u16 per 2 cells                       -> 1
#   32-bit:             2 v_min_u32 + 1 v_max3_u32 per cell x 2 steps                       -> 2/3
# Counter, 16-bit packing: one subtract + one signed dot product per 2 pair-votes.
VALU_LANE_INSTR = 256 * 64 * 2.4e9
COUNT_PEAK_K2 = VALU_LANE_INSTR
UPDATES_PER_INSTR = {"f16": 4.0 / 3.0, "u16": 1.0, "u32": 2.0 / 3.0}
PUBLISHED_TCELLS_32768 = 3.5  # reference README, Mojo on MI355X, end-to-end call


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).astype("<u4", copy=False).tobytes()).hexdigest()


def cpu_baseline(d: np.ndarray, budget_s: float = 15.0) -> dict:
    """Oracle tiled closure (T = 16) for a bounded number of pivot rounds of this matrix."""
    import oracle

    n = d.shape[0]
    T = 16
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    oracle.closure_tiled(d, T, threads=threads, max_rounds=0)  # p0 init only
    t_init = time.time() - t0
    t0 = time.time()
    oracle.closure_tiled(d, T, threads=threads, max_rounds=2)
    t_two = max(1e-3, time.time() - t0 - t_init)
    rounds = int(max(1, min((n + T - 1) // T, 2 * budget_s / t_two)))
    t0 = time.time()
    oracle.closure_tiled(d, T, threads=threads, max_rounds=rounds)
    el = max(1e-6, time.time() - t0 - t_init)
    cells = rounds * float(n) * n * T
    return {
        "value": round(cells / el / 1e12, 5),
        "unit": "TCells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {rounds} of {(n + T - 1) // T} pivot rounds (tile 16) of the same {n}-candidate "
                  f"matrix, {cells:.3e} cell updates in {el:.1f}s (p0 init excluded)",
        "host": host_cpu(),
        "small_configs": small_config_baselines(threads),
    }


def host_cpu() -> dict:
    """CPU model and logical CPU count of the host the baseline ran on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count()}


def small_config_baselines(threads: int) -> dict:
    """The 128 x 2000 election (serial and tiled oracle, whole closure) and the 4096 x 4096
    election (tiled oracle, whole closure), each beside the GPU closure of the same d
    through the drop-in compute_strongest_paths (host arrays in and out); best of 20 runs
    (128) or 3 (4096)."""
    import oracle
    from scalingelections_amd import api, synth

    out = {}
    for name, m, n, seed, serial in (("128x2000", 2000, 128, 1, True), ("4096x4096", 4096, 4096, 2, False)):
        R = synth.impartial_culture(m, n, seed)
        d = oracle.build_prefs(list(R))
        row = {}
        reps = 20 if n <= 256 else 3  # best of reps: sub-millisecond calls are noisy

        def best(fn):
            out, t_best = None, float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                out = fn()
                t_best = min(t_best, time.perf_counter() - t0)
            return out, t_best

        if serial:
            p_serial, el = best(lambda: oracle.closure_serial(d))
            row["cpu_serial_s"] = round(el, 6)
        p_tiled, el = best(lambda: oracle.closure_tiled(d, 16, threads=threads))
        row["cpu_tiled_s"] = round(el, 6)
        api.compute_strongest_paths(d, allow_gpu=True, tile_size=32)
        p_gpu, el = best(lambda: api.compute_strongest_paths(d, allow_gpu=True, tile_size=32))
        row["gpu_end_to_end_s"] = round(el, 6)
        row["best_of"] = reps
        row["cpu_tiled_threads"] = threads
        row["bit_exact"] = bool(np.array_equal(p_gpu, p_tiled) and (not serial or np.array_equal(p_gpu, p_serial)))
        out[name] = row
        log(f"small config {name}: {row}")
    return out


def recorded_traffic(mode: str, n: int):
    """HBM bytes per round of the independent-tile kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters need their own profiler runs)."""
    if mode != "f16" or n != 32768:
        return None
    path = os.path.join(ROOT, "profiles", "panel_indep_f16_closure32768_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_round:"):
            return float(line.split(":", 1)[1])
    return None


def closure_step_timed(L, _lib, d_ptr, p_ptr, n, flags):
    t = _lib.se_timing()
    _lib.check(L.se_closure_device(d_ptr, n, n, p_ptr, n, flags | _lib.SE_FLAG_TIME_KERNELS, None,
                                   ctypes.byref(t)))
    return t.as_dict()


def closure_by_size(sizes) -> dict:
    from scalingelections_amd import _lib, devmem, synth

    L = _lib.lib()
    out = {}
    for n in sizes:
        d = synth.synthesized_preferences(n, 3)
        dd = devmem.DeviceBuffer.from_array(d)
        pd = devmem.DeviceBuffer(d.nbytes)
        tm = closure_step_timed(L, _lib, dd.ptr, pd.ptr, n, 0)
        mode = "f16" if tm["f16_ordered"] else ("u16" if tm["width_bits"] == 16 else "u32")
        peak = VALU_LANE_INSTR * UPDATES_PER_INSTR[mode]
        reps = max(2, int(2e11 / n ** 3))
        devmem.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.se_closure_device(dd.ptr, n, n, pd.ptr, n, 0, None, None))
        devmem.device_synchronize()
        el = (time.perf_counter() - t0) / reps
        out[str(n)] = {"ms": round(el * 1e3, 3), "tcells_per_s": round(n ** 3 / el / 1e12, 3), "mode": mode,
                       "roofline_frac": round(n ** 3 / el / peak, 4)}
        log(f"closure n={n}: {out[str(n)]}")
    return out


def counter_section() -> dict:
    from scalingelections_amd import _lib, devmem, synth

    L = _lib.lib()
    m = n = 4096
    R = synth.impartial_culture(m, n, 2)
    pos = synth.positions_from_rankings(R)
    pd = devmem.DeviceBuffer.from_array(pos)
    dd = devmem.DeviceBuffer(n * n * 4)
    p_dev = devmem.DeviceBuffer(n * n * 4)
    ms = ctypes.c_double(0)
    for _ in range(2):
        _lib.check(L.se_count_device(pd.ptr, m, n, dd.ptr, n, None, ctypes.byref(ms)))
    best = float("inf")
    for _ in range(5):
        _lib.check(L.se_count_device(pd.ptr, m, n, dd.ptr, n, None, ctypes.byref(ms)))
        best = min(best, ms.value)
    nb = (n + 127) // 128
    # issued pair-votes: full 128 x 128 tiles I < J, and 3/4 of each diagonal tile (its
    # quadrant below the diagonal is written by complement, not counted)
    issued = m * ((nb * (nb - 1) // 2) * 128 * 128 + nb * 128 * 128 * 3 // 4)
    useful = m * n * (n - 1) / 2
    # whole election: count + closure + wins + ranking, device-resident
    wins = np.empty(n, np.uint32)
    rank = np.empty(n, np.uint32)
    devmem.device_synchronize()
    t0 = time.perf_counter()
    _lib.check(L.se_count_device(pd.ptr, m, n, dd.ptr, n, None, None))
    _lib.check(L.se_closure_device(dd.ptr, n, n, p_dev.ptr, n, 0, None, None))
    _lib.check(L.se_rank_device(p_dev.ptr, n, n, wins.ctypes.data, rank.ctypes.data, None))
    el = time.perf_counter() - t0
    d = dd.download(np.empty((n, n), np.uint32))
    fx_path = os.path.join(ROOT, "tests", "golden", "c2_4096x4096_seed2.json")
    parity = "no fixture"
    if os.path.exists(fx_path):
        fx = json.load(open(fx_path))
        p = p_dev.download(np.empty((n, n), np.uint32))
        ok = sha(d) == fx["d_sha256"] and sha(p) == fx["p_sha256"]
        parity = "bit-exact (d and p sha256 match the oracle fixture)" if ok else "MISMATCH"
    res = {
        "workload": "4096 impartial-culture ballots x 4096 candidates, seed 2, positions resident in HBM",
        "count_kernel_ms": round(best, 3),
        "pair_votes_per_s": round(useful / (best / 1e3), 1),
        "roofline": {"bound": "valu", "achieved": round(issued / (best / 1e3) / 1e12, 3),
                     "peak": round(COUNT_PEAK_K2 / 1e12, 2),
                     "unit": "T pair-votes/s (issued: I<J tiles and 3/4 of each diagonal tile, 2 voters per word)",
                     "frac": round(issued / (best / 1e3) / COUNT_PEAK_K2, 4),
                     "peak_basis": "one v_sub_u32 + one v_dot2c_i32_i16 per 2 pair-votes at the VALU issue rate"},
        "election_ms": round(el * 1e3, 2),
        "election_tcells_per_s": round(n ** 3 / el / 1e12, 3),
        "parity": parity,
    }
    log(f"counter {res}")
    return res


def batch_and_serving(do_serve: bool) -> tuple:
    from scalingelections_amd import batch, devmem, serve

    t0 = time.time()
    w = batch.mixed_size_workload(4096, seed=4)
    log(f"mixed batch generated in {time.time() - t0:.1f}s ({w.positions.size:.3e} positions)")
    db = batch.DeviceBatch(w)
    db.run()
    devmem.device_synchronize()
    t0 = time.perf_counter()
    tm = db.run()
    devmem.device_synchronize()
    t_batch = time.perf_counter() - t0
    Pb, Rb = db.raw()
    t0 = time.perf_counter()
    db.run(per_launch=True)
    devmem.device_synchronize()
    t_plain = time.perf_counter() - t0
    Pl, Rl = db.raw()
    po, ro = w.paths_offsets(), w.rank_offsets()
    mism = sum(1 for e in range(w.elections)
               if not (np.array_equal(Pb[po[e]:po[e + 1]], Pl[po[e]:po[e + 1]])
                       and np.array_equal(Rb[ro[e]:ro[e + 1]], Rl[ro[e]:ro[e + 1]])))
    bres = {
        "workload": "4096 elections, C log-uniform [8,512], m uniform [100,10000], Mallows phi U(0.5,1), seed 4; "
                    "positions resident in HBM",
        "batched_elections_per_s": round(w.elections / t_batch, 1),
        "per_launch_elections_per_s": round(w.elections / t_plain, 1),
        "speedup": round(t_plain / t_batch, 2),
        "batched_ms": round(t_batch * 1e3, 2), "count_ms": round(tm["count_ms"], 2),
        "closure_ms": round(tm["closure_ms"], 2),
        "pair_votes_per_s": round(tm["pair_votes"] / t_batch, 1),
        "closure_tcells_per_s": round(tm["cells"] / (tm["closure_ms"] / 1e3) / 1e12, 3),
        "mismatched_elections": mism,
    }
    log(f"batch {bres}")
    if not do_serve:
        return bres, None
    # ---- serving
    small_ref = {e: (Rb[ro[e]:ro[e + 1]].copy(), serve._sha(Pb[po[e]:po[e + 1]])) for e in range(w.elections)}
    del Pb, Pl, Rl
    t0 = time.time()
    large = batch.make_workload(np.array([8192, 8192], np.uint32), np.array([8192, 8192], np.uint32),
                                np.array([0.9, 1.0]), seed=5)
    log(f"large pool generated in {time.time() - t0:.1f}s")
    large_ref = {}
    t_large = []
    for e in range(large.elections):
        serve.solo_result(large, e)  # warm the library's scratch for this size
        rank_e, sha_e, secs = serve.solo_result(large, e, timed=True)
        large_ref[e] = (rank_e, sha_e)
        t_large.append(secs)
    t_large_s = float(np.median(t_large))
    small_rate = w.elections / t_batch
    large_workers = 2
    capacity = 1.0 / (0.99 / small_rate + 0.01 * t_large_s / large_workers)
    rate = 0.5 * capacity
    count = int(rate * 4.0)
    arr = serve.poisson_schedule(rate, count, seed=5)
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(count):
        if rng.random() < 0.01:
            reqs.append(serve.Request(k, float(arr[k]), large, int(rng.integers(large.elections))))
        else:
            reqs.append(serve.Request(k, float(arr[k]), w, int(rng.integers(w.elections))))
    srv = serve.Server(large_slots=large_workers, large_max_c=8192, large_max_m=8192)
    try:
        res = srv.run(reqs)
    finally:
        srv.close()
    bad = 0
    for r in res.requests:
        ref = large_ref[r.index] if r.large else small_ref[r.index]
        if not (np.array_equal(r.ranking, ref[0]) and r.paths_sha256 == ref[1]):
            bad += 1
    s = res.summary()
    sres = {
        "workload": "Poisson arrivals; sizes from the batch mix plus 1% 8192 x 8192 Mallows elections",
        "offered_req_per_s": round(rate, 1),
        "capacity_estimate_req_per_s": round(capacity, 1),
        "note": "offered load = 50% of the capacity estimated from the batched rate and the service time of "
                "one 8192 x 8192 election alone (upload, count, closure, ranking, download, digest) on each "
                "of 2 large workers; 20k req/s would exceed capacity and only measure queue growth",
        "completed_req_per_s": s["completed_per_s"],
        "requests": s["requests"], "batches": s["batches"], "mean_batch": s["mean_batch"],
        "latency_ms": {k: s[k] for k in ("all", "small", "large") if k in s},
        "large_solo_ms": round(t_large_s * 1e3, 1),
        "mismatched_requests": bad,
    }
    log(f"serving {sres}")
    return bres, sres


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the size sweep, counter, batch and serving")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (runtime copy kernel) and exit: the known-size "
                         "reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        from scalingelections_amd import devmem

        nbytes = args.copy_calibration << 20
        a, b = devmem.DeviceBuffer(nbytes), devmem.DeviceBuffer(nbytes)
        devmem._ok(devmem.hip().hipMemcpy(b.ptr, a.ptr, nbytes, devmem.hipMemcpyDeviceToDevice), "hipMemcpy D2D")
        devmem.device_synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs on one device; how the path would shard over "
              "several GPUs is described in DESIGN.md", file=sys.stderr)
        return 2

    from scalingelections_amd import _lib, api, devmem, synth

    info = api.device_info(0)
    n = args.n
    log(f"device {info['arch']} cus={info['compute_units']}; generating {n}x{n} preferences (seed {args.seed})")
    t0 = time.time()
    d = synth.synthesized_preferences(n, args.seed)
    log(f"generated in {time.time() - t0:.1f}s; uploading")
    d_dev = devmem.DeviceBuffer.from_array(d)
    p_dev = devmem.DeviceBuffer(d.nbytes)
    L = _lib.lib()

    for i in range(args.warmup):
        _lib.check(L.se_closure_device(d_dev.ptr, n, n, p_dev.ptr, n, 0, None, None))
        devmem.device_synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")

    timings = []
    devmem.device_synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = closure_step_timed(L, _lib, d_dev.ptr, p_dev.ptr, n, 0)  # synchronises on its last event
        timings.append(t)
        log(f"step {i + 1}/{args.steps}: {t['total_ms']:.1f} ms on device")
    devmem.device_synchronize()
    elapsed = time.perf_counter() - t_start

    ms_per_step = elapsed * 1e3 / max(1, args.steps)
    value = float(n) ** 3 / (ms_per_step / 1e3) / 1e12
    width = timings[-1]["width_bits"] if timings else 0
    mode = "u32" if width == 32 else ("f16" if timings and timings[-1]["f16_ordered"] else "u16")
    rounds_run = sum(t["rounds"] for t in timings)
    indep_ms = sum(t["indep_ms"] for t in timings)
    indep_cells = sum(t["indep_cells"] for t in timings)
    avg_round_ms = indep_ms / max(1, rounds_run)
    cells_per_round = indep_cells / max(1, rounds_run)
    achieved = cells_per_round / (avg_round_ms / 1e3) / 1e12 if avg_round_ms > 0 else 0.0
    peak = VALU_LANE_INSTR * UPDATES_PER_INSTR.get(mode, 1.0) / 1e12

    u32 = None
    if not args.no_extra:
        # the same matrix through the 32-bit closure (the 4 GB variant of this workload)
        log("32-bit closure of the same matrix (SE_FLAG_FORCE_U32): warmup")
        _lib.check(L.se_closure_device(d_dev.ptr, n, n, p_dev.ptr, n, _lib.SE_FLAG_FORCE_U32, None, None))
        devmem.device_synchronize()
        t0 = time.perf_counter()
        tm32 = closure_step_timed(L, _lib, d_dev.ptr, p_dev.ptr, n, _lib.SE_FLAG_FORCE_U32)
        devmem.device_synchronize()
        el32 = time.perf_counter() - t0
        fixture = os.path.join(ROOT, "tests", "golden", f"c3_{n}_seed{args.seed}.json")
        parity = "no fixture"
        if os.path.exists(fixture):
            got = p_dev.download(np.empty_like(d))
            parity = ("bit-exact (sha256 of p matches the oracle fixture)"
                      if sha(got) == json.load(open(fixture))["p_sha256"] else "MISMATCH")
            del got
        u32 = {"candidates": n, "ms": round(el32 * 1e3, 2), "tcells_per_s": round(float(n) ** 3 / el32 / 1e12, 3),
               "roofline_frac": round(float(n) ** 3 / el32 / (VALU_LANE_INSTR * UPDATES_PER_INSTR["u32"]), 4),
               "pivot_depth": tm32["pivot"], "width_bits": tm32["width_bits"], "parity": parity}
        log(f"u32 closure {u32}")
    del p_dev

    e2e = None
    if not args.no_e2e:
        log("end-to-end host call (upload + closure + download)")
        out = np.empty_like(d)
        t0 = time.perf_counter()
        rc = L.se_closure(d.ctypes.data, n, n, out.ctypes.data, 32, 0, None)
        e2e_s = time.perf_counter() - t0
        _lib.check(rc)
        fixture = os.path.join(ROOT, "tests", "golden", f"c3_{n}_seed{args.seed}.json")
        parity = "no fixture"
        if os.path.exists(fixture):
            want = json.load(open(fixture))["p_sha256"]
            parity = "bit-exact (sha256 of p matches the oracle fixture)" if sha(out) == want else "MISMATCH"
        e2e = {"seconds": round(e2e_s, 3), "tcells_per_s": round(float(n) ** 3 / e2e_s / 1e12, 3), "parity": parity}
        log(f"e2e {e2e}")
        del out

    extra = {}
    if u32 is not None:
        extra["closure_u32"] = u32
    if not args.no_extra:
        extra["closure_by_size"] = closure_by_size([4096, 8192, 16384])
        extra["counter"] = counter_section()
        extra["batch"], extra["serving"] = batch_and_serving(not args.no_serve)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, bounded rounds)")
        cpu = cpu_baseline(d)
        log(f"cpu {cpu}")

    last = timings[-1] if timings else {}
    line = {
        "metric": "TCells/s single election @32768 cand; elections/s batched vs per-launch",
        "value": round(value, 3),
        "unit": "TCells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED_TCELLS_32768, 3) if n == 32768 else None,
        "dtype": {"f16": "u16 (f16-ordered)", "u16": "u16", "u32": "u32"}.get(mode, f"u{width}"),
        "data": "synthetic",
        "config": {
            "workload": f"single election, {n} candidates, synthesized preferences (Binomial(65535, "
                        f"0.5+0.02z), 5% ties, seed {args.seed}); input resident in HBM",
            "candidates": n,
            "pivot_depth": last.get("pivot"),
            "padded_n": last.get("padded_n"),
            "rounds": last.get("rounds"),
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": (f"panel_indep_u16_kernel (each round's independent tiles with the next pivot's panel, "
                       f"one launch per round) + indep_u16_kernel (last round), {mode} instruction mix"
                       if width == 16 else "indep_u32_kernel (independent tiles, u32 instruction mix)"),
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(peak, 2),
            "unit": "TCells/s (cell updates; 2 integer ops each)",
            "frac": round(achieved / peak, 4) if peak else None,
            "traffic": recorded_traffic(mode, n),
            "traffic_note": "HBM bytes per round from rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a 1 GiB copy, "
                            "profiles/panel_indep_f16_closure32768_hbm_traffic.md; algorithmic 4.295e9 "
                            "(read + write the 16-bit matrix once)",
            "peak_basis": "256 CUs x 64 lanes x 2.4 GHz VALU issue x cell updates per lane-instruction "
                          "(f16-ordered 4/3, packed u16 1, u32 2/3)",
            "cells_per_round": cells_per_round,
            "avg_round_ms": round(avg_round_ms, 4),
            "rounds_timed": rounds_run,
            "note": "avg_round_ms: the round's launches (one in the look-ahead schedule used at this size, two "
                    "where the next pivot's strips run on their own), from HIP events on the library stream",
        },
        "phases_ms_per_step": {k: round(sum(t[k] for t in timings) / max(1, len(timings)), 2)
                               for k in ("total_ms", "init_ms", "panel_ms", "indep_ms", "output_ms")},
        "end_to_end": e2e,
        "vs_baseline_end_to_end": round(e2e["tcells_per_s"] / PUBLISHED_TCELLS_32768, 3) if (e2e and n == 32768) else None,
        **extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
