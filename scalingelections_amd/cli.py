"""Benchmark CLI with the reference's flags, defaults and report lines.

Reference: scaling_elections.py:299-565 (argparse flags :365-417, timing :320-357,
throughput n^3 / t :499 and :527, validation :538-542, election results :547-565).
Implementations benchmarked: the serial CPU closure (validation baseline), the tiled CPU
branch (``allow_gpu=False``) and the HIP GPU closure (``allow_gpu=True``, also with
``allow_tma=True``, which AMD GPUs ignore).  The reference's Numba entries are not
available here (Numba is not part of this package); the serial baseline is this
package's own serial CPU closure.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import api, builder, ranking


def format_time(elapsed_sec: float) -> str:
    """ms below one second, else seconds (scaling_elections.py:299-305)."""
    elapsed_ms = elapsed_sec * 1000
    if elapsed_ms < 1000:
        return f"{int(elapsed_ms)} ms"
    return f"{elapsed_sec:.2f} s"


def format_throughput(cells_per_sec: float) -> str:
    """T/G/M/K cells^3/s (scaling_elections.py:308-317)."""
    if cells_per_sec >= 1e12:
        return f"{cells_per_sec / 1e12:.1f} Tcells³/s"
    if cells_per_sec >= 1e9:
        return f"{cells_per_sec / 1e9:.1f} Gcells³/s"
    if cells_per_sec >= 1e6:
        return f"{cells_per_sec / 1e6:.1f} Mcells³/s"
    return f"{cells_per_sec / 1e3:.1f} Kcells³/s"


def benchmark_implementation(callback, preferences: np.ndarray, warmup: int, repeat: int):
    """(avg_time, result, success) over `repeat` timed calls after `warmup` calls
    (scaling_elections.py:320-357)."""
    for i in range(warmup):
        start_time = time.time()
        _ = callback(preferences)
        elapsed_time = time.time() - start_time
        if warmup > 1:
            print(f"  Warm-up {i+1}/{warmup}: {format_time(elapsed_time)}")
        else:
            print(f"  Warm-up: {format_time(elapsed_time)}")
    times = []
    result = None
    for _ in range(repeat):
        start_time = time.time()
        result = callback(preferences)
        times.append(time.time() - start_time)
    avg_time = sum(times) / len(times) if times else float("nan")
    return avg_time, result, True


def parse_args(argv=None):
    parser = argparse.ArgumentParser(description="Benchmark the Schulze method")
    parser.add_argument("--num-voters", type=int, default=0,
                        help="Number of voters in the population, 0 for random preference matrix")
    parser.add_argument("--num-candidates", type=int, default=256, help="Number of candidates in the election")
    parser.add_argument("--run-cpu", action="store_true", help="Run CPU implementations")
    parser.add_argument("--run-gpu", action="store_true", help="Run GPU implementation (HIP)")
    parser.add_argument("--no-serial", action="store_true", help="Skip serial baseline")
    parser.add_argument("--cpu-tile-size", type=int, default=16, help="CPU tile size for tiling optimization")
    parser.add_argument("--gpu-tile-size", type=int, default=32, help="GPU tile size for tiling optimization")
    parser.add_argument("--warmup", type=int, default=1, help="Number of warmup iterations (default: 1)")
    parser.add_argument("--repeat", type=int, default=1, help="Number of benchmark iterations (default: 1)")
    parser.add_argument("--seed", type=int, default=None,
                        help="Seed for the random inputs (the reference leaves them unseeded)")
    return parser.parse_args(argv)


def main(argv=None) -> int:
    args = parse_args(argv)
    cpu_tile_size = args.cpu_tile_size
    gpu_tile_size = args.gpu_tile_size
    num_voters = args.num_voters
    num_candidates = args.num_candidates
    run_serial = not args.no_serial
    rng = np.random.default_rng(args.seed)

    def gpu(x):
        return api.compute_strongest_paths(x, allow_gpu=True, allow_tma=False, tile_size=gpu_tile_size)

    def gpu_tma(x):
        return api.compute_strongest_paths(x, allow_gpu=True, allow_tma=True, tile_size=gpu_tile_size)

    def cpu(x):
        return api.compute_strongest_paths(x, allow_gpu=False, allow_tma=False, tile_size=cpu_tile_size)

    print("=== Schulze Voting Algorithm (Python) ===")
    print()
    try:
        api.log_gpus()
    except Exception as e:  # noqa: BLE001 -- reported, as in the reference
        print(f"✗ Could not detect GPU: {e}")

    print("Configuration:")
    voters_str = f"{num_voters:,}" if num_voters > 0 else "random"
    print(f"  Problem size: {num_candidates:,} candidates × {voters_str} voters")
    print(f"  CPU tile: {cpu_tile_size} × {cpu_tile_size}")
    if args.run_gpu:
        print(f"  GPU tile: {gpu_tile_size} × {gpu_tile_size}")
    print(f"  CPU threads: {os.cpu_count()}")
    print(f"  Warmup: {args.warmup}, Repeat: {args.repeat}")
    print()

    print("Generating preferences...")
    if num_voters == 0:
        preferences = rng.integers(0, num_candidates, (num_candidates, num_candidates)).astype(np.uint32)
    else:
        voter_rankings = [rng.permutation(num_candidates) for _ in range(num_voters)]
        preferences = builder.build_pairwise_preferences(voter_rankings)

    print()
    print("─── Benchmarking ───────────────────────────────────")
    print()

    serial_result = None
    if run_serial:
        print("→ Serial (C++)")
        try:
            avg_time, serial_result, _ = benchmark_implementation(
                api.compute_strongest_paths_serial, preferences, args.warmup, args.repeat)
            throughput = num_candidates ** 3 / avg_time
            if args.repeat > 1:
                print(f"  Run:     {format_time(avg_time)} (avg of {args.repeat}) │ {format_throughput(throughput)}")
            else:
                print(f"  Run:     {format_time(avg_time)} │ {format_throughput(throughput)}")
        except Exception as e:  # noqa: BLE001
            print(f"  ✗ Benchmark failed: {e}")
        print()

    mismatches = 0
    for name, wanted, callback in [
        ("Tiled CPU (C++ threads)", args.run_cpu, cpu),
        ("Tiled GPU (HIP)", args.run_gpu, gpu),
        ("Tiled GPU (HIP, allow_tma=True)", args.run_gpu, gpu_tma),
    ]:
        if not wanted:
            continue
        print(f"→ {name}")
        try:
            avg_time, result, _ = benchmark_implementation(callback, preferences, args.warmup, args.repeat)
            throughput = num_candidates ** 3 / avg_time
            if args.repeat > 1:
                print(f"  Run:     {format_time(avg_time)} (avg of {args.repeat}) │ {format_throughput(throughput)}")
            else:
                print(f"  Run:     {format_time(avg_time)} │ {format_throughput(throughput)}")
            if serial_result is not None:
                if np.array_equal(result, serial_result):
                    print("  ✓ Results validated")
                else:
                    mismatches += 1
                    print("  ✗ Results don't match baseline!")
        except Exception as e:  # noqa: BLE001
            print(f"  ✗ Benchmark failed: {e}")
        print()

    if serial_result is not None:
        result_for_winner = serial_result
    else:
        result_for_winner = api.compute_strongest_paths_serial(preferences)
    candidates = list(range(result_for_winner.shape[0]))
    winner, ranked = ranking.get_winner_and_ranking(candidates, result_for_winner)

    print("─── Election Results ───────────────────────────────")
    print()
    print(f"  Winner: Candidate #{winner}")
    if len(ranked) >= 5:
        print(f"  Top 5:  #{ranked[0]}, #{ranked[1]}, #{ranked[2]}, #{ranked[3]}, #{ranked[4]}")
    print()
    return 1 if mismatches else 0


if __name__ == "__main__":
    raise SystemExit(main())
