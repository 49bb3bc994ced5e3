"""Checks of the oracle that do not come from the reference run: closed forms,
textbook special cases, invariants, conservation, the exact Sod solution and
the reference's shipped Sod data (analysis/sod.ana, analysis/sodtest.dat).
A plausible slip in the restated algorithm (a dropped term, a wrong sign or
index) fails at least one of them."""
import os

import numpy as np
import pytest

import pyoracle as po
from boxzyhydro_amd import ics
from boxzyhydro_amd.params import HydroParams, box_params, sod_params
from conftest import ROOT
from sod_exact import sample

GOLD = os.path.join(ROOT, "tests", "golden")


# ------------------------------------------------------------------ limiter
def gmm(a, b, c):
    if a > 0 and b > 0 and c > 0:
        return min(a, b, c)
    if a < 0 and b < 0 and c < 0:
        return max(a, b, c)
    return 0.0


def test_limiter_closed_form():
    rng = np.random.default_rng(0)
    vals = rng.standard_normal((4000, 3))
    vals[:500] = np.round(vals[:500], 1)  # ties and zeros
    vals[500:600, 1] = vals[500:600, 0]
    for c0, c1, c2 in vals:
        want = gmm(1.25 * (c1 - c0), 0.5 * (c2 - c0), 1.25 * (c2 - c1))
        assert po.slope(c0, c1, c2) == want
    # monotone data take the smallest of the three candidates; extrema give 0
    assert po.slope(0.0, 1.0, 2.0) == 1.0  # min(1.25, 1.0, 1.25)
    assert po.slope(0.0, 1.0, 1.25) == 0.3125  # min(1.25, 0.625, 0.3125)
    assert po.slope(0.0, 1.0, 0.0) == 0.0
    assert po.slope(2.0, 1.0, 0.0) == -1.0


# --------------------------------------------------------------------- HLL
def physical_flux(d, v, p, e, idir):
    U = np.array([d, d * v[0], d * v[1], d * v[2], e + 0.5 * d * np.dot(v, v)])
    F = U * v[idir]
    F[1 + idir] += p
    F[4] += p * v[idir]
    return U, F


def test_hll_closed_form():
    rng = np.random.default_rng(1)
    g = 1.4
    for _ in range(300):
        idir = int(rng.integers(3))
        dl, dr = rng.uniform(0.1, 3, 2)
        pl, pr = rng.uniform(0.1, 3, 2)
        vl, vr = rng.uniform(-2, 2, (2, 3))
        el, er = pl / (g - 1), pr / (g - 1)
        L = [dl, *vl, pl, el, g]
        R = [dr, *vr, pr, er, g]
        F = po.face(idir, L, R)
        Ul, Fl = physical_flux(dl, vl, pl, el, idir)
        Ur, Fr = physical_flux(dr, vr, pr, er, idir)
        cl, cr = np.sqrt(g * pl / dl), np.sqrt(g * pr / dr)
        ap = max(0.0, vr[idir] + cr, vl[idir] + cl)
        am = max(0.0, -(vr[idir] - cr), -(vl[idir] - cl))
        want = (ap * Fl + am * Fr - ap * am * (Ur - Ul)) / (ap + am)  # Harten-Lax-van Leer, Davis speeds
        assert np.allclose(F, want, rtol=1e-13, atol=1e-13)
        # identical states give the physical flux; supersonic flow gives the upwind flux
        Fs = po.face(idir, L, L)
        assert np.allclose(Fs, Fl, rtol=1e-14, atol=1e-14)
    L = [1.0, 5.0, 0.0, 0.0, 1.0, 2.5, 1.4]
    R = [0.5, 4.0, 0.0, 0.0, 0.5, 1.25, 1.4]
    _, Fl = physical_flux(1.0, np.array([5.0, 0, 0]), 1.0, 2.5, 0)
    assert np.allclose(po.face(0, L, R), Fl, rtol=2e-16 * 4, atol=0)


# ----------------------------------------------------------------- constants
def test_units_and_muc_match_reference_print():
    """rgas of units.f90:10-17 as the reference printed it (sodtest.dat header)."""
    prm = sod_params()
    og = po.OracleGrid(prm)
    hdr = open(os.path.join(GOLD, "sodtest_profile.txt")).read().splitlines()[3].lstrip("#").split()
    rgas = po.lib().bxo_rgas(og._h)
    assert f"{rgas:.15E}" == f"{float(hdr[4]):.15E}"  # printed with 16 significant digits
    assert po.lib().bxo_muc(og._h) == 1.0 / (0.73 * 0.5 + 0.27 * 0.25 + 0.0 / 16.78)
    assert abs(prm.rgas - float(hdr[4])) < 1e-20


# -------------------------------------------------------------- exact Sod
def test_exact_solver_matches_reference_sod_ana():
    data = np.loadtxt(os.path.join(GOLD, "sod_ana.txt"))
    x = data[:, 0]
    ex = sample(x, 0.4, 0.0, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    bad = np.abs(ex[:, 0] - data[:, 1]) > 2e-4
    assert bad.sum() <= 4  # only samples that straddle a discontinuity
    assert np.allclose(ex[~bad, 1], data[~bad, 2], atol=2e-4)
    assert np.allclose(ex[~bad, 2], data[~bad, 3], atol=2e-4)


@pytest.fixture(scope="module")
def sod_run():
    prm = sod_params(endtime=0.2)
    ic = ics.sod_ic(prm)
    og = po.OracleGrid(prm, ic)
    b = ics.boundary_flags(prm).reshape(prm.ngrid) == 0
    vol = prm.dx * prm.dy * prm.dz
    tot0 = og.cons[b].sum(axis=0) * vol
    dts, times = og.run(100000)
    return prm, og, tot0, b, vol, times


def test_sod_profile_vs_exact(sod_run):
    prm, og, _, _, _, times = sod_run
    assert times[-1] == 0.2
    x, _, _ = ics.cell_centres(prm)
    row = og.cons.reshape(prm.nz, prm.ny, prm.nx, 5)[0, 2, 2:-2]
    xs = x[2:-2] + 0.5
    ex = sample(xs, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    l1 = np.sum(np.abs(row[:, 0] - ex[:, 0])) * prm.dx
    # the reference's own shipped run scores L1 = 8.7e-3 at dx=0.01 on [-1,1] (t=0.4)
    assert l1 < 4e-3, l1
    # post-shock plateau and contact speed
    assert abs(row[np.argmin(np.abs(xs - 0.75)), 0] - ex[np.argmin(np.abs(xs - 0.75)), 0]) < 3e-3


def test_sod_conserves_mass_momentum_energy(sod_run):
    prm, og, tot0, b, vol, _ = sod_run
    tot = og.cons[b].sum(axis=0) * vol
    assert abs(tot[0] - tot0[0]) < 1e-13 * tot0[0]
    assert abs(tot[4] - tot0[4]) < 1e-12 * tot0[4]
    # net x-momentum = pressure impulse through the (unmoved) ends: (p_L - p_R) * t * area
    area = (prm.ny - 4) * prm.nz * prm.dy * prm.dz
    assert abs(tot[1] - (1.0 - 0.1) * 0.2 * area) < 1e-10
    assert np.all(tot[2:4] == 0.0)


def test_y_uniform_state_stays_exactly_1d():
    prm = HydroParams(nx=40, ny=9, nz=7, dx=0.1, dy=0.1, dz=0.1, cfl=0.5, gammafix=1.4, endtime=1e30,
                      no_outflow_yl=False, no_outflow_yr=False, no_outflow_zl=False, no_outflow_zr=False)
    rng = np.random.default_rng(3)
    rho = np.broadcast_to(np.exp(rng.normal(0, 0.3, prm.nx)), prm.shape)
    p = np.broadcast_to(np.exp(rng.normal(0, 0.3, prm.nx)), prm.shape)
    vx = np.broadcast_to(rng.normal(0, 0.5, prm.nx), prm.shape)
    zero = np.zeros(prm.shape)
    ic = ics._assemble(prm, rho, vx, zero, zero, p)
    og = po.OracleGrid(prm, ic)
    og.run(6)
    c = og.cons.reshape(prm.nz, prm.ny, prm.nx, 5)[2:-2, 2:-2, 2:-2]
    assert np.all(c == c[0:1, 0:1])
    assert np.all(c[..., 2:4] == 0.0)


def test_uniform_state_is_steady():
    prm = box_params(10)
    n = prm.ngrid
    ic = ics._assemble(prm, np.ones(prm.shape), np.zeros(prm.shape), np.zeros(prm.shape),
                       np.zeros(prm.shape), np.ones(prm.shape))
    og = po.OracleGrid(prm, ic)
    og.run(4)
    b = ics.boundary_flags(prm).reshape(n) == 0
    assert np.all(og.cons[b, 0] == 1.0)
    assert np.all(og.cons[b, 1:4] == 0.0)
    assert np.all(og.cons[b, 4] == ic["cons"][b, 4])


def test_reference_setup_overlays_shipped_sodtest(golden):
    """The reference's own set-up (in.nml) at t=0.4 against its shipped
    snapshot analysis/sodtest.dat (an older code version: a visual overlay, so
    the bound is loose) and against the exact solution."""
    prof = np.array(golden["readme_sod_y"]["profile_ix4_iz1"]["rho"])
    ref = np.loadtxt(os.path.join(GOLD, "sodtest_profile.txt"))
    y = ref[:, 0]
    inner = slice(2, -2)
    l1_old = np.sum(np.abs(prof[inner] - ref[inner, 1])) * 0.01
    ex = sample(y[inner], 0.4, 0.0, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    l1_ex = np.sum(np.abs(prof[inner] - ex[:, 0])) * 0.01
    assert l1_old < 2e-2, l1_old
    assert l1_ex < 1.2e-2, l1_ex


def test_oracle_diagnostics_closed_forms():
    """The oracle's per-step diagnostics (conservation_diagnostic, utils.f90:1021-1060,
    run at main.f90:222-223 by the CPU baseline with diagnostics) on states
    whose sums have closed forms: a uniform state (mass, momenta, energies;
    centre of mass and angular momentum zero by symmetry) and a rigid rotation
    about z with a density ramp in x (angular momentum sum of rho omega r^2 dV,
    x centre of mass sum of x rho dV)."""
    prm = HydroParams(nx=14, ny=12, nz=10, dx=0.5, dy=0.25, dz=2.0)
    n, shp = prm.ngrid, prm.shape
    dv = prm.dx * prm.dy * prm.dz
    x, y, z = ics.cell_centres(prm)
    interior = ics.boundary_flags(prm) == 0
    V = interior.sum() * dv

    def grid(rho, m, E):
        ic = {"cons": np.zeros((n, 5)), "u": np.zeros((n, 3)), "p": np.ones(n), "adindx": np.full(n, 1.4)}
        ic["cons"][:, 0] = np.broadcast_to(rho, shp).reshape(n)
        for k in range(3):
            ic["cons"][:, 1 + k] = np.broadcast_to(m[k], shp).reshape(n)
            ic["u"][:, k] = ic["cons"][:, 1 + k] / ic["cons"][:, 0]
        ic["cons"][:, 4] = np.broadcast_to(E, shp).reshape(n)
        og = po.OracleGrid(prm, ic)
        return og.diagnostics()

    d = grid(2.0, (1.0, -0.5, 0.25), 3.0)
    # mass, momx, momy, momz, amom, etot, ekin, xcom, ycom, zcom, em, ef
    want = [2 * V, 1 * V, -0.5 * V, 0.25 * V, 0.0, 3 * V, (1 + 0.25 + 0.0625) / 4 * V, 0.0, 0.0, 0.0, 2 * V, 0.0]
    for k, (a, b) in enumerate(zip(d, want)):
        assert a == pytest.approx(b, rel=1e-12, abs=1e-12 * V), k

    om = 0.3
    X, Y = np.broadcast_to(x[None, None, :], shp), np.broadcast_to(y[None, :, None], shp)
    rho = 1.0 + 0.1 * X
    d = grid(rho, (-om * Y * rho, om * X * rho, 0.0 * X), 1.0)
    r2 = (X * X + Y * Y)[interior]
    assert d[4] == pytest.approx(float(np.sum(om * rho[interior] * r2)) * dv, rel=1e-12)
    assert d[7] == pytest.approx(float(np.sum(X[interior] * rho[interior])) * dv, rel=1e-12)
    assert d[0] == pytest.approx(float(np.sum(rho[interior])) * dv, rel=1e-12)
