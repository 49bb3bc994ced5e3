"""Host-side mirror of the reference's hydro-step interface on the GPU.

The reference's hot path is a set of argument-less Fortran subroutines acting
on module globals (grid_commons.f90:7-36) and called from main.f90:211-555.
``HydroGrid`` keeps that shape: it owns the device copy of those globals
(cons, cons_old, u, p, adindx, qq) and the scalars ``dt`` and ``time``, and
exposes one method per subroutine with the reference's name and argument
meaning:

=====================  ==========================  ==========================
method                 reference                   C-ABI entry point
=====================  ==========================  ==========================
``state()``            state.f90:6-89              bx_state
``velocity()``         velocity.f90:4-64           bx_velocity
``courant()``          courant.f90:4-39 (sets dt)  bx_courant
``source(tphase,it)``  source.f90:5-205 (uses dt)  bx_source
``set_ghost_cells()``  utils.f90:305-480           bx_set_ghost_cells
``flux(avg)``          flux.f90:6-235 (uses dt)    bx_flux
``avisc()``            avisc.f90:4-26              bx_avisc
``copy_old()``         main.f90:327-333            bx_copy_old
``run(...)``           main.f90:196-555            bx_step
=====================  ==========================  ==========================

Host arrays are the reference's Fortran layout: ``cons`` (ngrid, 5) etc.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Tuple

import numpy as np

from . import native as nv
from .params import HydroParams


def make_params(prm: HydroParams, precision: int = nv.BX_PREC_FP64) -> nv.BxParams:
    P = nv.BxParams()
    P.nx, P.ny, P.nz = prm.nx, prm.ny, prm.nz
    P.dx, P.dy, P.dz = prm.dx, prm.dy, prm.dz
    P.gammafix = prm.gammafix
    P.small_rho, P.small_eps, P.vlimit, P.cfl = prm.small_rho, prm.small_eps, prm.vlimit, prm.cfl
    P.xabun, P.yabun, P.zabun, P.mu_z = prm.xabun, prm.yabun, prm.zabun, prm.mu_z
    for i, v in enumerate((prm.avmagx, prm.avmagy, prm.avmagz)):
        P.avmag[i] = v
    for i, v in enumerate((prm.no_outflow_xl, prm.no_outflow_xr, prm.no_outflow_yl,
                           prm.no_outflow_yr, prm.no_outflow_zl, prm.no_outflow_zr)):
        P.no_outflow[i] = int(bool(v))
    P.precision = precision
    P.fluxangmom = int(bool(prm.fluxangmom))
    P.yoffset = prm.yoffset
    return P


def make_eos_params(prm: HydroParams) -> nv.BxEosParams:
    """bx_eos_params of ``prm``: the &eos_input keys, the rest at the
    reference's parameters.f90 values."""
    E = nv.BxEosParams()
    nv.lib().bx_eos_defaults(ctypes.byref(E))
    E.H2STAT, E.NEOS_RHO = prm.H2STAT, prm.NEOS_RHO
    E.dTk_eos, E.tk_bgrnd, E.brot, E.vib, E.ac, E.bc = (prm.dTk_eos, prm.tk_bgrnd, prm.brot, prm.vib,
                                                         prm.ac, prm.bc)
    return E


def eos_tables(prm: HydroParams) -> Dict[str, np.ndarray]:
    """Host-only: the tables calc_eos_table (eos.f90:127-520) leaves for
    ``prm``; 2-D tables have shape (NEOS_RHO, NEOS_T)."""
    E = make_eos_params(prm)
    nt = ctypes.c_int32()
    nv.check(nv.lib().bx_eos_size(ctypes.byref(E), ctypes.byref(nt)), "bx_eos_size")
    nt, nr = nt.value, prm.NEOS_RHO
    out = {"scalars": np.empty(2), "rho_table": np.empty(nr), "deng_eos_array": np.empty(nr),
           "tk_table": np.empty(nt)}
    names2 = ("eng_table", "gamma_table", "muc_table", "eng_table2", "tk_table2", "muc_table2", "gamma_table2")
    for k in names2:
        out[k] = np.empty((nr, nt))
    order = ("scalars", "rho_table", "deng_eos_array", "tk_table") + names2
    nv.check(nv.lib().bx_eos_tables(ctypes.byref(make_params(prm)), ctypes.byref(E),
                                    *[nv.dptr(out[k]) for k in order]), "bx_eos_tables")
    out["log_rho_eos_low"], out["drho_eos"] = float(out["scalars"][0]), float(out["scalars"][1])
    out["NEOS_T"], out["NEOS_RHO"] = nt, nr
    return out


class HydroGrid:
    """Device-resident state of one grid plus the reference's step scalars."""

    def __init__(self, prm: HydroParams, ic: Optional[Dict[str, np.ndarray]] = None,
                 device: int = 0, precision: int = nv.BX_PREC_FP64):
        self.prm = prm
        self._h = ctypes.c_void_p()
        nv.check(nv.lib().bx_create(ctypes.byref(make_params(prm, precision)), device,
                                    ctypes.byref(self._h)), "bx_create")
        ng = ctypes.c_int64()
        gh = ctypes.c_int64()
        nv.check(nv.lib().bx_ngrid(self._h, ctypes.byref(ng), ctypes.byref(gh)), "bx_ngrid")
        self.ngrid, self.nghost = ng.value, gh.value
        # main.f90:194-203 driver scalars
        self.dt = 0.0
        self.time = prm.starttime
        self.dtold = 0.0
        self.step_count = 0
        if prm.H2STAT >= 0:  # main.f90:70-71: initialize_eos + calc_eos_table
            nv.check(nv.lib().bx_eos_init(self._h, ctypes.byref(make_eos_params(prm))), "bx_eos_init")
        if ic is not None:
            self.upload(ic)

    def close(self) -> None:
        if self._h:
            nv.lib().bx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ transfers
    def upload(self, ic: Dict[str, np.ndarray]) -> None:
        n = self.ngrid
        arrs = []
        for key, k in (("cons", 5), ("u", 3), ("p", 1), ("adindx", 1)):
            a = ic.get(key)
            if a is not None:
                a = np.ascontiguousarray(a, np.float64)
                if a.size != n * k:
                    raise ValueError(f"{key}: expected {n * k} values, got {a.size}")
            arrs.append(a)
        nv.check(nv.lib().bx_upload(self._h, *[nv.dptr(a) for a in arrs]), "bx_upload")

    def download(self) -> Dict[str, np.ndarray]:
        n = self.ngrid
        out = {"cons": np.empty((n, 5)), "u": np.empty((n, 3)), "p": np.empty(n), "adindx": np.empty(n)}
        nv.check(nv.lib().bx_download(self._h, nv.dptr(out["cons"]), nv.dptr(out["u"]),
                                      nv.dptr(out["p"]), nv.dptr(out["adindx"])), "bx_download")
        return out

    # -------------------------------------------------------------- gravity
    def set_gravity(self, flags: int, phi: Optional[np.ndarray] = None, object_mass: float = 1.0,
                    object_radius: float = 0.625, initial: bool = True) -> None:
        """Gravity from a given potential ``phi`` (ngrid): BX_GRAV_POTENTIAL
        (calc_gforce), BX_GRAV_EXTERNAL (external_grav's point mass),
        BX_GRAV_COM_SEQUENTIAL (mass moments in cell order, bit-identical to a
        one-thread reference; otherwise a tree order).  ``initial`` makes the
        calls of main.f90:146-154 that precede the step loop."""
        self.grav_flags = flags
        nv.check(nv.lib().bx_gravity_config(self._h, flags, object_mass, object_radius), "bx_gravity_config")
        if not flags:
            return
        if phi is not None:
            phi = np.ascontiguousarray(phi, np.float64)
            if phi.size != self.ngrid:
                raise ValueError(f"phi: expected {self.ngrid} values, got {phi.size}")
        nv.check(nv.lib().bx_set_phi(self._h, nv.dptr(phi)), "bx_set_phi")
        if initial:
            if flags & nv.BX_GRAV_POTENTIAL:
                self.calc_gforce()
            if flags & nv.BX_GRAV_EXTERNAL:
                self.external_grav()

    def calc_gforce(self) -> None:
        """utils.f90:549-571 on the device."""
        nv.check(nv.lib().bx_calc_gforce(self._h), "bx_calc_gforce")

    def external_grav(self) -> np.ndarray:
        """gravity.f90:1477-1549; returns the mass moments (x, y, z)."""
        com = np.empty(3)
        nv.check(nv.lib().bx_external_grav(self._h, nv.dptr(com)), "bx_external_grav")
        return com

    def download_gforce(self) -> np.ndarray:
        out = np.empty((self.ngrid, 3))
        nv.check(nv.lib().bx_download_gforce(self._h, nv.dptr(out)), "bx_download_gforce")
        return out

    def download_muc(self) -> np.ndarray:
        """muc_array (grid_commons.f90:12) as the last state() left it."""
        out = np.empty(self.ngrid)
        nv.check(nv.lib().bx_download_muc(self._h, nv.dptr(out)), "bx_download_muc")
        return out

    def download_cons_old(self) -> np.ndarray:
        out = np.empty((self.ngrid, 5))
        nv.check(nv.lib().bx_download_cons_old(self._h, nv.dptr(out)), "bx_download_cons_old")
        return out

    # ------------------------------------------------ reference subroutines
    def state(self) -> None:
        nv.check(nv.lib().bx_state(self._h), "bx_state")

    def velocity(self) -> float:
        maxv = ctypes.c_double()
        nv.check(nv.lib().bx_velocity(self._h, ctypes.byref(maxv)), "bx_velocity")
        return maxv.value

    def courant(self) -> float:
        dt = ctypes.c_double()
        nv.check(nv.lib().bx_courant(self._h, ctypes.byref(dt)), "bx_courant")
        self.dt = dt.value
        return self.dt

    def source(self, tphase: float, siter: int) -> None:
        nv.check(nv.lib().bx_source(self._h, tphase, siter, self.dt), "bx_source")

    def set_ghost_cells(self) -> None:
        nv.check(nv.lib().bx_set_ghost_cells(self._h), "bx_set_ghost_cells")

    def flux(self, avg: int) -> None:
        nv.check(nv.lib().bx_flux(self._h, avg, self.dt), "bx_flux")

    def copy_old(self) -> None:
        nv.check(nv.lib().bx_copy_old(self._h), "bx_copy_old")

    def avisc(self) -> np.ndarray:
        qq = np.empty((self.ngrid, 3))
        nv.check(nv.lib().bx_avisc(self._h, nv.dptr(qq)), "bx_avisc")
        return qq

    def sync(self) -> None:
        nv.check(nv.lib().bx_sync(self._h), "bx_sync")

    # ------------------------------------------------------------ main loop
    def run(self, nsteps: int, endtime: Optional[float] = None, mode: int = nv.BX_STEP_FUSED,
            record_dt: bool = True) -> np.ndarray:
        """Advance up to ``nsteps`` steps of main.f90:211-555 (or to endtime).

        Returns the full dt of every step taken."""
        if endtime is None:
            endtime = self.prm.endtime
        t = ctypes.c_double(self.time)
        dto = ctypes.c_double(self.dtold)
        st = ctypes.c_int32(self.step_count)
        taken = ctypes.c_int32(0)
        dts = np.zeros(max(nsteps, 1)) if record_dt else None
        nv.check(nv.lib().bx_step(self._h, mode, nsteps, endtime, ctypes.byref(t), ctypes.byref(dto),
                                  ctypes.byref(st), nv.dptr(dts), ctypes.byref(taken)), "bx_step")
        self.time, self.dtold, self.step_count = t.value, dto.value, st.value
        return dts[:taken.value].copy() if record_dt else np.zeros(0)

    def tune(self, y_chunk: int, z_chunk: int, occupancy: int = 0) -> None:
        """Work split of the fused step's stage kernels (bx_tune): rows per
        wave of the y-kernel, planes per wave of the x/z kernel or of the
        tiled stage kernel, and the variant (0, 1, 4: two kernels per stage;
        10, 11: one tiled kernel).  Results unchanged."""
        nv.check(nv.lib().bx_tune(self._h, y_chunk, z_chunk, occupancy), "bx_tune")

    def profile(self, enable: int = -1):
        """Event timing of the fused stage kernels: enable=1 start, 0 stop,
        -1 read.  Returns ([y1_ms, xz1_ms, y2_ms, xz2_ms] summed, steps)."""
        ms = (ctypes.c_double * 4)()
        steps = ctypes.c_int64()
        nv.check(nv.lib().bx_profile(self._h, enable, ms, ctypes.byref(steps)), "bx_profile")
        return list(ms), steps.value

    def write_celldump(self, path: str, step: int, ascii: bool = False) -> None:
        """write_files(step): state(), then the celldump file `path`."""
        nv.check(nv.lib().bx_write_celldump(self._h, os.fsencode(path), self.time, step, int(ascii)),
                 "bx_write_celldump")

    def read_celldump(self, path: str) -> int:
        """read_hydro(): restart from the celldump `path`; returns its step
        and sets the driver time (starttime) from it."""
        t = ctypes.c_double()
        s = ctypes.c_int32()
        nv.check(nv.lib().bx_read_celldump(self._h, os.fsencode(path), ctypes.byref(t), ctypes.byref(s)),
                 "bx_read_celldump")
        self.time = t.value
        return s.value

    def graph_step_count(self) -> int:
        """Steps taken from graph replays (BX_STEP_GRAPH) so far."""
        c = ctypes.c_int64()
        nv.check(nv.lib().bx_graph_step_count(self._h, ctypes.byref(c)), "bx_graph_step_count")
        return c.value

    def step_counts(self):
        """(steps taken with the fused kernels, steps taken with the exact
        sequence) since the grid was created (bx_step_counts)."""
        f, e = ctypes.c_int64(), ctypes.c_int64()
        nv.check(nv.lib().bx_step_counts(self._h, ctypes.byref(f), ctypes.byref(e)), "bx_step_counts")
        return f.value, e.value

    def fallback_count(self) -> int:
        c = ctypes.c_int64()
        nv.check(nv.lib().bx_fallback_count(self._h, ctypes.byref(c)), "bx_fallback_count")
        return c.value


def celldump_encode(prm: HydroParams, path: str, time: float, step: int, cons, u, p, ascii: bool = False,
                    muc=None) -> None:
    """Host-only celldump writer on reference-layout arrays already updated by
    state(); ``muc`` is muc_array (None: the fixed-gamma muc everywhere)."""
    c = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (cons, u, p, muc)]
    nv.check(nv.lib().bx_celldump_encode(ctypes.byref(make_params(prm)), os.fsencode(path), time, step, int(ascii),
                                         *[nv.dptr(a) for a in c]), "bx_celldump_encode")


def celldump_decode(prm: HydroParams, path: str):
    """Host-only celldump reader: (time, step, arrays as read_hydro holds them
    before its final state() call)."""
    n = prm.ngrid
    out = {"cons": np.empty((n, 5)), "u": np.empty((n, 3)), "p": np.empty(n), "adindx": np.empty(n)}
    t = ctypes.c_double()
    s = ctypes.c_int32()
    nv.check(nv.lib().bx_celldump_decode(ctypes.byref(make_params(prm)), os.fsencode(path), ctypes.byref(t),
                                         ctypes.byref(s), nv.dptr(out["cons"]), nv.dptr(out["u"]), nv.dptr(out["p"]),
                                         nv.dptr(out["adindx"])), "bx_celldump_decode")
    return t.value, s.value, out


def step_sequence(grid, nsteps: int, endtime: float) -> Tuple[list, list]:
    """The main.f90:211-555 call sequence written out through the per-subroutine
    methods -- what a driver keeping the reference's calls does.  Works on any
    object with the reference's subroutine names (HydroGrid or the oracle)."""
    dts, times = [], []
    f1 = 1.0
    while grid.time < endtime and grid.step_count < nsteps:
        grid.step_count += 1
        grid.state()
        dt = grid.courant()
        dt = dt * f1
        if grid.step_count > 1 and grid.dtold > 0.0:
            dt = min(dt, grid.dtold * 1.1)
        grid.dtold = dt
        if dt + grid.time > endtime:
            dt = endtime - grid.time
        grid.dt = dt * 0.5
        grid.source(grid.time, 0)
        grid.set_ghost_cells()
        grid.state()
        grid.velocity()
        grid.copy_old()
        grid.dt = grid.dt * 2.0
        grid.flux(0)
        grid.set_ghost_cells()
        grid.state()
        grid.velocity()
        grid.flux(1)
        grid.set_ghost_cells()
        grid.state()
        grid.velocity()
        grid.dt = 0.5 * grid.dt
        grid.source(grid.time + grid.dt, 1)
        grid.set_ghost_cells()
        grid.state()
        grid.velocity()
        grid.dt = grid.dt * 2.0
        grid.time = grid.time + grid.dt
        dts.append(grid.dt)
        times.append(grid.time)
    return dts, times
