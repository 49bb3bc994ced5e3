"""Autotuner for the fused step's work split.

Times a few steps of each candidate (y rows per wave of the y-direction
kernel, z planes per wave of the x/z kernel, variant of the x/z kernel) on the grid it is given and keeps the fastest.  Results of the step do
not depend on the choice, so tuning can run on the production grid itself; the
state it leaves behind is stepped forward by the timing steps, so tune on a
copy (or re-upload the initial condition afterwards).
"""
from __future__ import annotations

import itertools
import time
from typing import Iterable, List, Tuple

from . import native as nv

DEFAULT_CANDIDATES = list(itertools.product((16, 32, 64), (16, 32, 64), (0, 1, 4)))


def tune_step(grid, candidates: Iterable[Tuple[int, int, int]] = DEFAULT_CANDIDATES, steps: int = 3,
              warmup: int = 1) -> Tuple[Tuple[int, int, int], List[Tuple[Tuple[int, int, int], float]]]:
    """Return (best candidate, [(candidate, seconds per step), ...])."""
    results = []
    for cand in candidates:
        grid.tune(*cand)
        grid.run(warmup, mode=nv.BX_STEP_FUSED, endtime=1e300)
        grid.sync()
        t0 = time.perf_counter()
        grid.run(steps, mode=nv.BX_STEP_FUSED, endtime=1e300)
        grid.sync()
        results.append((cand, (time.perf_counter() - t0) / steps))
    best = min(results, key=lambda r: r[1])[0]
    grid.tune(*best)
    return best, results
