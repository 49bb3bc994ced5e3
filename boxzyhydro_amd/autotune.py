"""Autotuner for the fused step's work split.

Times a few steps of each candidate (y rows per wave of the y-direction
kernel, z planes per wave of the x/z kernel or of the tiled stage kernel, the
stage variant) on the grid it is given and keeps the fastest.  Results of the step do
not depend on the choice, so tuning can run on the production grid itself; the
state it leaves behind is stepped forward by the timing steps, so tune on a
copy (or re-upload the initial condition afterwards).
"""
from __future__ import annotations

import itertools
import time
from typing import Iterable, List, Tuple

from . import native as nv

# (y rows, z planes, variant): the two-kernel stages (variants 0, 1, 4) and
# the one-kernel tiled stages (variants 10, 11; y rows unused)
DEFAULT_CANDIDATES = (list(itertools.product((16, 32, 64), (16, 32, 64), (0, 1, 4))) +
                      list(itertools.product((64,), (16, 32, 64), (10, 11))))


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


BATCH_CANDIDATES = [(16, 1), (16, 2), (8, 1), (8, 2), (4, 2), (4, 4)]


def tune_batch(batch, candidates: Iterable[Tuple[int, int]] = BATCH_CANDIDATES, reps: int = 2,
               graph: bool = True) -> Tuple[Tuple[int, int], List[Tuple[Tuple[int, int], float]]]:
    """Work split of the batched-tube launch (waves per workgroup, workgroups
    per compute unit): each candidate runs the whole batch ``reps`` times
    from its set-up (device time of the launch, best of reps); the fastest is
    kept.  Results do not depend on the choice."""
    results = []
    for cand in candidates:
        batch.tune(*cand)
        best = float("inf")
        for _ in range(reps):
            batch.run(graph=graph)
            best = min(best, batch.sync())
        results.append((cand, best * 1e-3))
    best = min(results, key=lambda r: r[1])[0]
    batch.tune(*best)
    return best, results


# (members per gang, gangs): each of the costliest tubes over several workgroups
GANG_CANDIDATES = [(0, 0), (4, 16), (4, 32), (4, 48), (6, 24), (6, 32), (8, 16)]


def tune_gangs(batch, candidates: Iterable[Tuple[int, int]] = GANG_CANDIDATES, reps: int = 2,
               graph: bool = True) -> Tuple[Tuple[int, int], List[Tuple[Tuple[int, int], float]]]:
    """Gangs of the batched-tube launch (bx_batch_gang): with the work split
    already chosen, each candidate runs the whole batch ``reps`` times (the
    first run also measures each tube's steps, which pick the gang tubes); the
    fastest is kept.  Candidates the device cannot hold (members x gangs above
    its compute units) are skipped.  Results do not depend on the choice."""
    batch.gang(0, 0)
    batch.run(graph=graph)
    batch.sync()  # the measured steps of every tube
    results = []
    for cand in candidates:
        try:
            batch.gang(*cand)
        except nv.BoxzyError:
            continue
        best = float("inf")
        for _ in range(reps):
            batch.run(graph=graph)
            best = min(best, batch.sync())
        results.append((cand, best * 1e-3))
    best = min(results, key=lambda r: r[1])[0]
    batch.gang(*best)
    return best, results


# the 4- and 8-wave builds of the batched-grid kernel hold about 210 registers
# per lane (8 waves per CU); the 16-wave build is held to 128 and spills
BOX_CANDIDATES = [(16, 1), (8, 1), (4, 1), (4, 2)]


def tune_boxes(batch, candidates: Iterable[Tuple[int, int]] = BOX_CANDIDATES, reps: int = 2,
               graph: bool = True) -> Tuple[Tuple[int, int], List[Tuple[Tuple[int, int], float]]]:
    """Work split of the batched-grid launch (waves per workgroup, workgroups
    per compute unit), timed like tune_batch.  Results do not depend on the
    choice."""
    return tune_batch(batch, candidates, reps, graph)
