"""MI355X-native (gfx950) Schulze engine: the reference's Python entry points over a HIP
library reached through a C ABI.

Drop-ins (reference file:line):
  compute_strongest_paths, log_gpus, reduce   -- scaling_elections.cu:863-1038 (pybind11)
  build_pairwise_preferences                  -- scaling_elections.py:51-88
  get_winner_and_ranking                      -- scaling_elections.py:274-296
"""
from .api import (  # noqa: F401
    compute_strongest_paths,
    compute_strongest_paths_serial,
    device_count,
    device_info,
    interrupt,
    log_gpus,
    reduce,
    strongest_paths_timed,
)
from .builder import build_pairwise_preferences, count_positions  # noqa: F401
from .ranking import get_winner_and_ranking, wins_and_order  # noqa: F401

__all__ = [
    "compute_strongest_paths",
    "log_gpus",
    "reduce",
    "build_pairwise_preferences",
    "get_winner_and_ranking",
]
