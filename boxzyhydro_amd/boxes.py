"""Batched grids on one GPU (the ``bx_boxes_*`` entry points of
include/boxzy.h).

Many independent grids of uneven size -- 3-D boxes, or flat-z slabs with
nz < 6 -- run in one persistent launch, optionally replayed from a captured
hipGraph.  Each grid has its own parameters and endtime and returns what the
reference's main loop (main.f90:211-555) returns for it in a process of its
own, bit for bit: the fixed-gamma EOS (H2STAT=-1), the plain flux
(fluxangmom=0) and the force of a zero potential, in FP64.  A grid steps as
``bx_step(BX_STEP_FUSED)`` steps it: fused steps on a regular state, the
exact call sequence where required.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import native as nv
from .hydro import make_params
from .params import HydroParams

Grid = Dict[str, np.ndarray]

STATUS_BADDT = 4  # a NaN, zero or infinite dt stopped the grid before that step


class GridBatch:
    """Many independent grids in one persistent launch on one GPU.

    ``grids`` is a sequence of (HydroParams, set-up) with the set-up in the
    reference's Fortran layout: ``cons`` (ngrid,5), ``u`` (ngrid,3), ``p``
    (ngrid,) -- what ``HydroGrid`` uploads, minus ``adindx`` (set by the first
    state call)."""

    def __init__(self, grids: Sequence[Tuple[HydroParams, Grid]], endtimes: Optional[Sequence[float]] = None,
                 device: int = 0):
        self.grids = list(grids)
        self.n = len(self.grids)
        if any(prm.starttime != 0.0 for prm, _ in self.grids):
            raise ValueError("batched grids start at time 0 (starttime = 0)")
        self.ngrid = np.array([prm.ngrid for prm, _ in self.grids], np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.ngrid)])
        if endtimes is None:
            endtimes = [prm.endtime for prm, _ in self.grids]
        self.endtimes = np.ascontiguousarray(endtimes, np.float64).reshape(-1)
        if len(self.endtimes) != self.n:
            raise ValueError(f"{len(self.endtimes)} endtimes for {self.n} grids")
        for i, (prm, g) in enumerate(self.grids):
            for key, k in (("cons", 5), ("u", 3), ("p", 1)):
                if np.asarray(g[key]).size != prm.ngrid * k:
                    raise ValueError(f"grid {i}: {key} has {np.asarray(g[key]).size} values, expected {prm.ngrid * k}")
        cons = np.ascontiguousarray(np.concatenate([np.asarray(g["cons"], np.float64).reshape(-1, 5)
                                                    for _, g in self.grids]))
        u = np.ascontiguousarray(np.concatenate([np.asarray(g["u"], np.float64).reshape(-1, 3) for _, g in self.grids]))
        p = np.ascontiguousarray(np.concatenate([np.asarray(g["p"], np.float64).reshape(-1) for _, g in self.grids]))
        arr = (nv.BxParams * self.n)(*[make_params(prm) for prm, _ in self.grids])
        self._h = ctypes.c_void_p()
        nv.check(nv.lib().bx_boxes_create(self.n, arr, nv.dptr(self.endtimes), nv.dptr(cons), nv.dptr(u),
                                           nv.dptr(p), device, ctypes.byref(self._h)), "bx_boxes_create")
        self.maxlog = 0

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nv.lib().bx_boxes_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def cells(self) -> Tuple[int, int]:
        """(cells of all grids, their interior cells)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        nv.check(nv.lib().bx_boxes_cells(self._h, ctypes.byref(a), ctypes.byref(b)), "bx_boxes_cells")
        return a.value, b.value

    def order(self) -> np.ndarray:
        o = np.empty(self.n, np.int32)
        nv.check(nv.lib().bx_boxes_order(self._h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                 "bx_boxes_order")
        return o

    def tune(self, waves_per_block: int, blocks_per_cu: int) -> None:
        nv.check(nv.lib().bx_boxes_tune(self._h, waves_per_block, blocks_per_cu), "bx_boxes_tune")

    def log_dt(self, maxlog: int) -> None:
        """Keep the first `maxlog` dt of each grid in the next runs (0: no log)."""
        self.maxlog = 0  # the library turns the log off before it allocates the new one
        nv.check(nv.lib().bx_boxes_log_dt(self._h, maxlog), "bx_boxes_log_dt")
        self.maxlog = maxlog

    def run(self, maxsteps: int = 1_000_000, graph: bool = True) -> None:
        """Enqueue one run of every grid from its set-up (asynchronous)."""
        nv.check(nv.lib().bx_boxes_run(self._h, maxsteps, 1 if graph else 0), "bx_boxes_run")

    def sync(self) -> float:
        """Wait for the run; returns its device time in ms (HIP events)."""
        ms = ctypes.c_double()
        nv.check(nv.lib().bx_boxes_sync(self._h, ctypes.byref(ms)), "bx_boxes_sync")
        return ms.value

    def result(self) -> List[Dict[str, object]]:
        tot = int(self.off[-1])
        cons = np.empty((tot, 5))
        u = np.empty((tot, 3))
        p = np.empty(tot)
        t = np.empty(self.n)
        steps = np.empty(self.n, np.int32)
        status = np.empty(self.n, np.uint32)
        nv.check(nv.lib().bx_boxes_result(self._h, nv.dptr(cons), nv.dptr(u), nv.dptr(p), nv.dptr(t),
                                          steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          status.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                 "bx_boxes_result")
        out = []
        for i in range(self.n):
            a, b = int(self.off[i]), int(self.off[i + 1])
            out.append({"cons": cons[a:b], "u": u[a:b], "p": p[a:b], "time": float(t[i]),
                        "steps": int(steps[i]), "status": int(status[i])})
        return out

    def dts(self) -> List[np.ndarray]:
        """dt of each step of each grid (up to the log length)."""
        if not self.maxlog:
            raise nv.BoxzyError("dt logging is off: call log_dt(maxlog) before run()")
        buf = np.empty((self.n, self.maxlog))
        nv.check(nv.lib().bx_boxes_dts(self._h, nv.dptr(buf)), "bx_boxes_dts")
        res = self.result()
        return [buf[i, :min(res[i]["steps"], self.maxlog)].copy() for i in range(self.n)]
