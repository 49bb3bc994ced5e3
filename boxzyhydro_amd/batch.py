"""Batched shock tubes on one GPU (the ``bx_batch_*`` entry points of
include/boxzy.h).

A tube is a grid with flat z (nz < 6), outflow on every side and the same
set-up in every (y,z) row apart from the frame cells; it is passed as that
x-row (``cons`` (nx,5), ``u`` (nx,3), ``p`` (nx,)).  All tubes run in one
persistent launch, optionally replayed from a captured hipGraph, and each
returns what the reference's main loop (main.f90:211-555) returns for the
full grid: every row of that grid equals the returned row, with ghost cells
holding their donors' values.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import native as nv
from .hydro import make_params
from .ics import boundary_flags
from .params import HydroParams

Rows = Dict[str, np.ndarray]

STATUS_LOWRHO = 1  # the low-density EOS branch (state.f90:52-56) was met
STATUS_CLAMP = 2   # the vlimit clamp (velocity.f90:33-50) was met
STATUS_BADDT = 4   # a non-finite dt stopped the tube


def tube_row_of(prm: HydroParams, ic: Dict[str, np.ndarray]) -> Rows:
    """The row of a 3-D set-up that is a tube; raises ValueError otherwise.

    Every non-frame cell must equal its x-row's value in row (iy=3, iz=1),
    frame cells must hold (small_rho, 0, 0, 0, small_eps), and u and p must be
    the row's everywhere (bitwise)."""
    shp = prm.shape
    cons = np.asarray(ic["cons"], np.float64).reshape(shp + (5,))
    u = np.asarray(ic["u"], np.float64).reshape(shp + (3,))
    p = np.asarray(ic["p"], np.float64).reshape(shp)
    row = {"cons": cons[0, 2].copy(), "u": u[0, 2].copy(), "p": p[0, 2].copy()}
    frame = boundary_flags(prm) == 1

    def same(a, b):
        return np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))

    if not (same(u, np.broadcast_to(row["u"], u.shape)) and same(p, np.broadcast_to(row["p"], p.shape))):
        raise ValueError("u or p differ between rows: not a tube")
    want = np.broadcast_to(row["cons"], cons.shape).copy()
    want[frame] = (prm.small_rho, 0.0, 0.0, 0.0, prm.small_eps)
    if not same(cons, want):
        raise ValueError("cons differ between rows (outside the frame): not a tube")
    return row


def expand_row(prm: HydroParams, row: Rows) -> Dict[str, np.ndarray]:
    """The full grid a returned row stands for, in the reference's Fortran
    layout: cons (ngrid,5), u (ngrid,3), p (ngrid,)."""
    shp = prm.shape
    return {"cons": np.ascontiguousarray(np.broadcast_to(row["cons"], shp + (5,)).reshape(-1, 5)),
            "u": np.ascontiguousarray(np.broadcast_to(row["u"], shp + (3,)).reshape(-1, 3)),
            "p": np.ascontiguousarray(np.broadcast_to(row["p"], shp).reshape(-1))}


def grid_of_row(prm: HydroParams, row: Rows) -> Dict[str, np.ndarray]:
    """The 3-D set-up a tube row stands for (reference layout, frame cells
    rho=small_rho, m=0, E=small_eps, adindx=gammafix)."""
    g = expand_row(prm, row)
    frame = boundary_flags(prm).reshape(-1) == 1
    g["cons"][frame] = (prm.small_rho, 0.0, 0.0, 0.0, prm.small_eps)
    g["adindx"] = np.full(prm.ngrid, prm.gammafix)
    return g


def _concat(tubes: Sequence[Tuple[HydroParams, Rows]]):
    for i, (prm, r) in enumerate(tubes):
        for key, k in (("cons", 5), ("u", 3), ("p", 1)):
            if np.asarray(r[key]).size != prm.nx * k:
                raise ValueError(f"tube {i}: {key} has {np.asarray(r[key]).size} values, expected {prm.nx * k} "
                                 f"(one x-row)")
    cons = np.ascontiguousarray(np.concatenate([r["cons"] for _, r in tubes]), np.float64)
    u = np.ascontiguousarray(np.concatenate([r["u"] for _, r in tubes]), np.float64)
    p = np.ascontiguousarray(np.concatenate([r["p"] for _, r in tubes]), np.float64)
    arr = (nv.BxParams * len(tubes))(*[make_params(prm) for prm, _ in tubes])
    return arr, cons, u, p


def check_tubes(tubes: Sequence[Tuple[HydroParams, Rows]]) -> None:
    """Host-side check of the row form (bx_batch_check); raises BoxzyError
    naming the first tube that does not meet it."""
    arr, cons, u, _ = _concat(tubes)
    bad = ctypes.c_int32(-1)
    nv.check(nv.lib().bx_batch_check(len(tubes), arr, nv.dptr(cons), nv.dptr(u), ctypes.byref(bad)),
             "bx_batch_check")


class TubeBatch:
    """Many independent tubes in one persistent launch on one GPU."""

    def __init__(self, tubes: Sequence[Tuple[HydroParams, Rows]], endtimes=None, device: int = 0):
        self.tubes = list(tubes)
        self.n = len(self.tubes)
        self.nx = np.array([prm.nx for prm, _ in self.tubes], np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.nx)])
        if endtimes is None:
            endtimes = [prm.endtime for prm, _ in self.tubes]
        self.endtimes = np.ascontiguousarray(endtimes, np.float64).reshape(-1)
        if len(self.endtimes) != self.n:
            raise ValueError(f"{len(self.endtimes)} endtimes for {self.n} tubes")
        arr, cons, u, p = _concat(self.tubes)
        self._h = ctypes.c_void_p()
        nv.check(nv.lib().bx_batch_create(self.n, arr, nv.dptr(self.endtimes), nv.dptr(cons), nv.dptr(u),
                                           nv.dptr(p), device, ctypes.byref(self._h)), "bx_batch_create")
        self.maxlog = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nv.lib().bx_batch_destroy(h)
            self._h = None

    def cells(self) -> Tuple[int, int]:
        """(cells of all rows, interior cells of the full grids)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        nv.check(nv.lib().bx_batch_cells(self._h, ctypes.byref(a), ctypes.byref(b)), "bx_batch_cells")
        return a.value, b.value

    def row_interior_cells(self) -> int:
        return int(np.sum(self.nx - 4))

    def order(self) -> np.ndarray:
        o = np.empty(self.n, np.int32)
        nv.check(nv.lib().bx_batch_order(self._h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "bx_batch_order")
        return o

    def tune(self, waves_per_block: int, blocks_per_cu: int) -> None:
        nv.check(nv.lib().bx_batch_tune(self._h, waves_per_block, blocks_per_cu), "bx_batch_tune")

    def gang(self, members: int, ntubes: int) -> None:
        """Split each of the ``ntubes`` costliest tubes over ``members``
        workgroups (bx_batch_gang; 0 turns it off).  Results do not change."""
        nv.check(nv.lib().bx_batch_gang(self._h, members, ntubes), "bx_batch_gang")

    def gang_fallbacks(self) -> int:
        """Runs repeated without gangs because a gang's members were not
        resident together."""
        n = ctypes.c_int32()
        nv.check(nv.lib().bx_batch_gang_fallbacks(self._h, ctypes.byref(n)), "bx_batch_gang_fallbacks")
        return n.value

    def log_dt(self, maxlog: int) -> None:
        nv.check(nv.lib().bx_batch_log_dt(self._h, maxlog), "bx_batch_log_dt")
        self.maxlog = maxlog

    def run(self, maxsteps: int = 1_000_000, graph: bool = True) -> None:
        """Enqueue one run of every tube from its set-up (asynchronous)."""
        nv.check(nv.lib().bx_batch_run(self._h, maxsteps, 1 if graph else 0), "bx_batch_run")

    def sync(self) -> float:
        """Wait for the run; returns its device time in ms (HIP events)."""
        ms = ctypes.c_double()
        nv.check(nv.lib().bx_batch_sync(self._h, ctypes.byref(ms)), "bx_batch_sync")
        return ms.value

    def result(self) -> List[Dict[str, object]]:
        tot = int(self.off[-1])
        cons = np.empty((tot, 5))
        u = np.empty((tot, 3))
        p = np.empty(tot)
        t = np.empty(self.n)
        steps = np.empty(self.n, np.int32)
        status = np.empty(self.n, np.uint32)
        nv.check(nv.lib().bx_batch_result(self._h, nv.dptr(cons), nv.dptr(u), nv.dptr(p), nv.dptr(t),
                                          steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          status.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                 "bx_batch_result")
        out = []
        for i in range(self.n):
            a, b = int(self.off[i]), int(self.off[i + 1])
            out.append({"cons": cons[a:b], "u": u[a:b], "p": p[a:b], "time": float(t[i]),
                        "steps": int(steps[i]), "status": int(status[i])})
        return out

    def dts(self) -> List[np.ndarray]:
        """dt of each step of each tube (up to the log length)."""
        if not self.maxlog:
            raise nv.BoxzyError("dt logging is off: call log_dt(maxlog) before run()")
        buf = np.empty((self.n, self.maxlog))
        nv.check(nv.lib().bx_batch_dts(self._h, nv.dptr(buf)), "bx_batch_dts")
        res = self.result()
        return [buf[i, :min(res[i]["steps"], self.maxlog)].copy() for i in range(self.n)]
