"""Drop-in for the reference's ``scaling_elections`` module and script.

``import scaling_elections`` exposes the reference's native-module API
(compute_strongest_paths, log_gpus, reduce -- scaling_elections.cu:991-1038) and its
Python helpers (build_pairwise_preferences, get_winner_and_ranking --
scaling_elections.py:51-88, :274-296), all backed by the gfx950 HIP library.
``python scaling_elections.py [flags]`` runs the reference's benchmark CLI
(scaling_elections.py:361-565) with the same flags and defaults.
"""
from scalingelections_amd import (  # noqa: F401
    build_pairwise_preferences,
    compute_strongest_paths,
    get_winner_and_ranking,
    log_gpus,
    reduce,
)
from scalingelections_amd.cli import benchmark_implementation, format_throughput, format_time  # noqa: F401

if __name__ == "__main__":
    from scalingelections_amd.cli import main

    raise SystemExit(main())
