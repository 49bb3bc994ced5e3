"""Initial conditions for the hydro-step workloads.

Every generator returns the reference's global arrays in Fortran order
(grid_commons.f90:7-36): ``cons`` (ngrid, 5), ``u`` (ngrid, 3), ``p`` (ngrid,),
``adindx`` (ngrid,), with cell id ``ix + (iy-1)nx + (iz-1)nx ny``
(utils.f90:191-194) -- i.e. a C-order (nz, ny, nx) grid flattened.

They follow the reference's set-up convention (init_conditions.f90:26-77):
every cell, ghosts included, gets the IC formula with ``E = p/(gamma-1) +
kinetic`` and ``adindx = gammafix``; then the outermost ghost layer
(boundary=1) gets ``rho = small_rho``, ``m = 0``, ``E = small_eps`` while
``p`` and ``u`` keep their IC values.  These frame cells enter the first
step's Courant maximum through the low-density branch of state.f90:52-56.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

from .params import HydroParams, box_params


def boundary_flags(prm: HydroParams) -> np.ndarray:
    """Ghost flags of init_grid.f90:81-118 as a (nz, ny, nx) int8 array."""
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    ix = np.arange(1, nx + 1)
    iy = np.arange(1, ny + 1)
    iz = np.arange(1, nz + 1)
    f1x = (ix == 1) | (ix == nx)
    f2x = (ix == 2) | (ix == nx - 1)
    f1y = (iy == 1) | (iy == ny)
    f2y = (iy == 2) | (iy == ny - 1)
    if nz >= 6:
        f1z = (iz == 1) | (iz == nz)
        f2z = (iz == 2) | (iz == nz - 1)
    else:
        f1z = np.zeros(nz, bool)
        f2z = np.zeros(nz, bool)
    f1 = f1z[:, None, None] | f1y[None, :, None] | f1x[None, None, :]
    f2 = f2z[:, None, None] | f2y[None, :, None] | f2x[None, None, :]
    out = np.zeros((nz, ny, nx), np.int8)
    out[f2] = 2
    out[f1] = 1
    return out


def cell_centres(prm: HydroParams) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Centred coordinates x = dx*(ix-(nx+1)/2) (init_grid.f90:85-93)."""
    x = prm.dx * (np.arange(1, prm.nx + 1, dtype=np.float64) - (prm.nx + 1.0) / 2.0)
    y = prm.dy * (np.arange(1, prm.ny + 1, dtype=np.float64) - (prm.ny + 1.0) / 2.0)
    z = prm.dz * (np.arange(1, prm.nz + 1, dtype=np.float64) - (prm.nz + 1.0) / 2.0)
    return x, y, z


def _assemble(prm: HydroParams, rho, vx, vy, vz, p) -> Dict[str, np.ndarray]:
    """IC arrays from primitive fields (each (nz,ny,nx)), reference convention."""
    n = prm.ngrid
    gam = prm.gammafix
    rho = np.ascontiguousarray(rho, np.float64).reshape(n)
    v = np.stack([np.asarray(a, np.float64).reshape(n) for a in (vx, vy, vz)], axis=1)
    p = np.ascontiguousarray(p, np.float64).reshape(n).copy()
    cons = np.empty((n, 5), np.float64)
    cons[:, 0] = rho
    cons[:, 1] = rho * v[:, 0]
    cons[:, 2] = rho * v[:, 1]
    cons[:, 3] = rho * v[:, 2]
    cons[:, 4] = p / (gam - 1.0) + rho * 0.5 * ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    frame = boundary_flags(prm).reshape(n) == 1
    cons[frame, 0] = prm.small_rho
    cons[frame, 1:4] = 0.0
    cons[frame, 4] = prm.small_eps
    adindx = np.full(n, gam, np.float64)
    return {"cons": cons, "u": np.ascontiguousarray(v), "p": p, "adindx": adindx}


def sod_ic(prm: HydroParams, axis: int = 0, x0: float = 0.0) -> Dict[str, np.ndarray]:
    """Sod tube (init_conditions.f90:31-53) split at centred coordinate ``x0``
    along ``axis`` (0=x, 1=y, 2=z): rho=1, p=1 below; rho=0.125, p=0.1 above."""
    x, y, z = cell_centres(prm)
    coord = [x[None, None, :], y[None, :, None], z[:, None, None]][axis]
    left = np.broadcast_to(coord < x0, prm.shape)
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 1.0 / 10.0)
    zero = np.zeros(prm.shape)
    return _assemble(prm, rho, zero, zero, zero, p)


def _smooth_unit(field: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian smoothing (reflecting edges), rescaled to zero mean and
    unit standard deviation so that the stated amplitudes survive smoothing."""
    from scipy.ndimage import gaussian_filter

    s = gaussian_filter(field, sigma=sigma, mode="reflect")
    s = s - s.mean()
    sd = s.std()
    return s / sd if sd > 0 else s


def box_ic(prm: HydroParams, seed: int = 0, sigma_cells: float = 4.0,
           rho_sigma: float = 0.1, p_sigma: float = 0.05,
           mach_rms: float = 0.5) -> Dict[str, np.ndarray]:
    """Seeded random 3-D box.

    * ``rho = exp(rho_sigma * g1)``, ``p = 1 + p_sigma * g2`` with g1, g2 white
      Gaussian fields smoothed by a Gaussian of ``sigma_cells`` cells and
      renormalised to unit variance;
    * velocities: independent white Gaussian components scaled so that the rms
      of |v| is ``mach_rms`` times c = sqrt(gamma <p>/<rho>).
    The generator is numpy PCG64 seeded with ``seed``; the whole grid (ghosts
    included) is drawn, then the reference frame convention is applied.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    shp = prm.shape
    g1 = _smooth_unit(rng.standard_normal(shp), sigma_cells)
    g2 = _smooth_unit(rng.standard_normal(shp), sigma_cells)
    rho = np.exp(rho_sigma * g1)
    p = 1.0 + p_sigma * g2
    c = np.sqrt(prm.gammafix * p.mean() / rho.mean())
    vel = rng.standard_normal((3,) + shp)
    vel *= (mach_rms * c) / np.sqrt(np.mean(np.sum(vel * vel, axis=0)))
    return _assemble(prm, rho, vel[0], vel[1], vel[2], p)


def rotating_box_ic(prm: HydroParams, seed: int = 0, omega: float = 0.2, sigma_cells: float = 2.0,
                    mach_rms: float = 0.3, yoffset: float = 0.0) -> Dict[str, np.ndarray]:
    """``box_ic`` plus solid rotation about the z axis through y = -yoffset,
    u += omega (-(y + yoffset), x, 0): the set-up for flux_ang
    (fluxangmom=.true.), which fluxes r- and angular momentum about that axis."""
    base = box_ic(prm, seed=seed, sigma_cells=sigma_cells, mach_rms=mach_rms)
    x, y, _ = cell_centres(prm)
    X = np.broadcast_to(x[None, None, :], prm.shape).reshape(-1)
    Y = np.broadcast_to(y[None, :, None], prm.shape).reshape(-1) + yoffset
    rho = base["cons"][:, 0].copy()
    interior = boundary_flags(prm).reshape(-1) != 1
    rho[~interior] = 1.0
    v = base["u"].copy()
    v[:, 0] -= omega * Y
    v[:, 1] += omega * X
    if prm.nz == 1:
        v[:, 2] = 0.0
    shp = prm.shape
    return _assemble(prm, rho.reshape(shp), v[:, 0].reshape(shp), v[:, 1].reshape(shp), v[:, 2].reshape(shp),
                     base["p"].reshape(shp))


def h2_cloud_ic(prm: HydroParams, seed: int = 0, sigma_cells: float = 2.0, rho0: float = 0.1,
                rho_sigma: float = 1.0, t_lo: float = 20.0, t_hi: float = 4000.0,
                mach_rms: float = 0.5) -> Dict[str, np.ndarray]:
    """Seeded molecular-cloud box for the tabulated H2 EOS (H2STAT >= 0), in
    code units (units.f90): ``rho = rho0 exp(rho_sigma g1)`` and a temperature
    log-uniform-ish in [t_lo, t_hi] from ``g2`` (g1, g2 smoothed unit Gaussian
    fields); ``p = rho R T / muc`` (eos.f90:144 muc) and ``E = p/(gammafix-1)
    + ekin``, so the first state() call finds T through the tables.  rms
    Mach ``mach_rms`` against sqrt(1.4 <p>/<rho>)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shp = prm.shape
    g1 = _smooth_unit(rng.standard_normal(shp), sigma_cells)
    g2 = _smooth_unit(rng.standard_normal(shp), sigma_cells)
    rho = rho0 * np.exp(rho_sigma * g1)
    lt = 0.5 * (np.log(t_lo) + np.log(t_hi)) + 0.25 * (np.log(t_hi) - np.log(t_lo)) * g2
    tk = np.exp(np.clip(lt, np.log(t_lo), np.log(t_hi)))
    p = rho * prm.rgas * tk / prm.muc
    c = np.sqrt(1.4 * p.mean() / rho.mean())
    vel = rng.standard_normal((3,) + shp)
    vel *= (mach_rms * c) / np.sqrt(np.mean(np.sum(vel * vel, axis=0)))
    if prm.nz == 1:
        vel[2] = 0.0
    return _assemble(prm, rho, vel[0], vel[1], vel[2], p)


def h2_disk_ic(prm: HydroParams, seed: int = 0, omega: float = 0.05, **kw) -> Dict[str, np.ndarray]:
    """A single-layer (nz=1) rotating disk for the tabulated EOS's disk branch
    (state.f90:61-69, which needs a nonzero kinetic energy): the cloud of
    ``h2_cloud_ic`` plus solid rotation u = omega (-y, x, 0)."""
    if prm.nz != 1:
        raise ValueError("h2_disk_ic: nz must be 1")
    ic = h2_cloud_ic(prm, seed=seed, **kw)
    x, y, _ = cell_centres(prm)
    X = np.broadcast_to(x[None, None, :], prm.shape).reshape(-1)
    Y = np.broadcast_to(y[None, :, None], prm.shape).reshape(-1)
    rho = ic["cons"][:, 0]
    p = ic["p"]
    vx = ic["u"][:, 0] - omega * Y
    vy = ic["u"][:, 1] + omega * X
    zero = np.zeros_like(vx)
    return _assemble(prm, rho.reshape(prm.shape), vx.reshape(prm.shape), vy.reshape(prm.shape),
                     zero.reshape(prm.shape), p.reshape(prm.shape))


def _riemann_draws(n_problems: int, seed: int, lmin: int, lmax: int):
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = np.rint(np.exp(rng.uniform(np.log(lmin), np.log(lmax), n_problems))).astype(int)
    lr = np.exp(rng.uniform(np.log(0.1), np.log(10.0), (n_problems, 4)))
    vv = rng.uniform(-1.0, 1.0, (n_problems, 2))
    return lens, lr, vv


def _riemann_params(L: int, endtime: float, gamma: float, vlimit: float) -> HydroParams:
    d = 1.0 / L
    return HydroParams(nx=L + 4, ny=8, nz=4, dx=d, dy=d, dz=d,
                       no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False,
                       no_outflow_yr=False, no_outflow_zl=False, no_outflow_zr=False,
                       small_rho=1e-12, small_eps=1e-40, vlimit=vlimit, endtime=endtime,
                       dtout=1e30, cfl=0.5, gammafix=gamma)


def riemann_batch(n_problems: int = 512, seed: int = 0, lmin: int = 64, lmax: int = 4096,
                  endtime: float = 0.1, gamma: float = 1.4,
                  vlimit: float = 1e3) -> List[Tuple[HydroParams, Dict[str, np.ndarray]]]:
    """Independent 1-D Riemann problems of uneven length, as full 3-D grids.

    Lengths are log-uniform in [lmin, lmax] cells (each L x 4 x 4 interior,
    dx = 1/L so every tube spans [0,1]); left/right rho and p log-uniform in
    [0.1, 10], velocities uniform in [-1, 1]; diaphragm at the tube centre
    (x=0.5 of [0,1]).  ``vlimit`` is raised above the reference default so the
    speed clamp (velocity.f90:33-50) does not act on strong rarefactions.
    """
    lens, lr, vv = _riemann_draws(n_problems, seed, lmin, lmax)
    out = []
    for i in range(n_problems):
        prm = _riemann_params(int(lens[i]), endtime, gamma, vlimit)
        x, _, _ = cell_centres(prm)
        left = np.broadcast_to((x < 0.0)[None, None, :], prm.shape)
        rho = np.where(left, lr[i, 0], lr[i, 1])
        p = np.where(left, lr[i, 2], lr[i, 3])
        vx = np.where(left, vv[i, 0], vv[i, 1])
        zero = np.zeros(prm.shape)
        out.append((prm, _assemble(prm, rho, vx, zero, zero, p)))
    return out


def tube_row(prm: HydroParams, rho, vx, p) -> Dict[str, np.ndarray]:
    """The x-row of a shock-tube set-up (every (y,z) row of the grid equal),
    by the reference convention: E = p/(gamma-1) + kinetic, and the frame
    cells (ix = 1, nx) rho=small_rho, m=0, E=small_eps with the row's u, p."""
    nx = prm.nx
    rho = np.asarray(rho, np.float64).reshape(nx)
    vx = np.asarray(vx, np.float64).reshape(nx)
    p = np.asarray(p, np.float64).reshape(nx).copy()
    gam = prm.gammafix
    zero = np.zeros(nx)
    cons = np.empty((nx, 5), np.float64)
    cons[:, 0] = rho
    cons[:, 1] = rho * vx
    cons[:, 2] = rho * zero
    cons[:, 3] = rho * zero
    cons[:, 4] = p / (gam - 1.0) + rho * 0.5 * ((vx * vx + zero * zero) + zero * zero)
    for c in (0, nx - 1):
        cons[c] = (prm.small_rho, 0.0, 0.0, 0.0, prm.small_eps)
    u = np.ascontiguousarray(np.stack([vx, zero, zero], axis=1))
    return {"cons": cons, "u": u, "p": p}


def riemann_tubes(n_problems: int = 512, seed: int = 0, lmin: int = 64, lmax: int = 4096,
                  endtime: float = 0.1, gamma: float = 1.4,
                  vlimit: float = 1e3) -> List[Tuple[HydroParams, Dict[str, np.ndarray]]]:
    """The problems of :func:`riemann_batch` as tube rows (bit-identical to
    the rows of the 3-D set-ups), without building the 3-D grids."""
    lens, lr, vv = _riemann_draws(n_problems, seed, lmin, lmax)
    out = []
    for i in range(n_problems):
        prm = _riemann_params(int(lens[i]), endtime, gamma, vlimit)
        x, _, _ = cell_centres(prm)
        left = x < 0.0
        out.append((prm, tube_row(prm, np.where(left, lr[i, 0], lr[i, 1]),
                                  np.where(left, vv[i, 0], vv[i, 1]), np.where(left, lr[i, 2], lr[i, 3]))))
    return out


def box_workload(n_interior: int, seed: int = 0):
    prm = box_params(n_interior)
    return prm, box_ic(prm, seed=seed)


def write_ic_file(path: str, ic: Dict[str, np.ndarray]) -> None:
    """Stream-binary IC file read by the reference driver: cons u p adindx,
    then phi when ``ic`` has one."""
    with open(path, "wb") as f:
        for k in ("cons", "u", "p", "adindx", "phi"):
            if k == "phi" and k not in ic:
                continue
            f.write(np.ascontiguousarray(ic[k], np.float64).tobytes())


def to_planes(cons: np.ndarray, prm: HydroParams) -> np.ndarray:
    """Fortran cons(5,ngrid) -> SoA planes (5, nz, ny, nx)."""
    return np.ascontiguousarray(cons.T.reshape((5,) + prm.shape))


def from_planes(planes: np.ndarray) -> np.ndarray:
    """SoA planes (k, nz, ny, nx) -> Fortran-order (ngrid, k)."""
    k = planes.shape[0]
    return np.ascontiguousarray(planes.reshape(k, -1).T)
