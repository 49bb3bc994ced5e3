"""Grids with reflecting walls (no_outflow_*, utils.f90:338-403) against the
reference's own runs (tests/golden/walls.json, one thread).

Every call site of the step makes two set_ghost_cells calls in a row.  Where
each ghost's donors over that pair resolve to interior cells (every wall set
of a flat-z grid; 3-D boxes with walls on at most one low side) the fused step
fills the ghosts from the pair's composed gather table.  A 3-D box walled in
y and z (the reference's shipped in.nml walls every side) also has 16 corner
ghosts whose pair leaves another ghost's value from before it; the fused step
writes those after its ghost fill as the exact sequence leaves them.  Every
case takes the fused kernels after the first step.  All bit for bit: dt
values exactly, arrays through SHA-256 of their float64 bytes."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from cases import case_inputs, corner_blob_ic, ic_digest, sha

HERE = os.path.dirname(os.path.abspath(__file__))
FUSED = ["box64_xwalls", "box40_ywalls_zr", "flat196x100_allwalls", "box64_allwalls"]


@pytest.fixture(scope="module")
def walls():
    with open(os.path.join(HERE, "golden", "walls.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", FUSED)
def test_oracle_matches_reference_walls(walls, name):
    rec = walls[name]
    prm, ic = case_inputs(rec)
    assert ic_digest(ic) == rec["ic_sha256"], "input generator changed"
    og = po.OracleGrid(prm, ic)
    dts, _ = og.run(rec["steps"])
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(getattr(og, k)) == rec["final_sha256"][k], k


@pytest.mark.gpu
@pytest.mark.parametrize("name", FUSED)
@pytest.mark.parametrize("mode", ["fused", "graph"])
def test_gpu_walls_match_reference(walls, name, mode):
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = walls[name]
    prm, ic = case_inputs(rec)
    print(f"case {name} {mode}", flush=True)
    g = HydroGrid(prm, ic)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_FUSED if mode == "fused" else nv.BX_STEP_GRAPH)
    assert list(dts) == rec["dt"]
    out = g.download()
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    fused, exact = g.step_counts()
    print(f"  fused steps {fused}, exact steps {exact}, fallbacks {g.fallback_count()}", flush=True)
    assert fused + exact == rec["steps"]
    assert exact == 1 and g.fallback_count() == 0  # the first step from the set-up only
    if mode == "graph":
        assert g.graph_step_count() > 0
    g.close()


@pytest.mark.gpu
def test_gpu_walls_static_step_matches_fused():
    """bx_step_static fills walled ghosts from the same composed table."""
    import ctypes

    import torch

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = box_params(32)
    prm.no_outflow_xl = prm.no_outflow_xr = True
    ic = ics.box_ic(prm, seed=31, sigma_cells=2.0)
    g = HydroGrid(prm, ic)
    g.run(1, endtime=1e30, mode=nv.BX_STEP_FUSED)
    planes = []
    for _ in range(2):
        buf = torch.empty(5 * g.ngrid, dtype=torch.float64, device="cuda")
        nv.check(nv.lib().bx_cons_device(g._h, ctypes.c_void_p(buf.data_ptr()), 0), "bx_cons_device")
        planes.append(buf)
        if len(planes) == 1:
            dt = g.run(1, endtime=1e30, mode=nv.BX_STEP_FUSED)
    out = torch.empty_like(planes[0])
    nv.check(nv.lib().bx_step_static(g._h, ctypes.c_void_p(planes[0].data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                     float(dt[0]), None, None, None), "bx_step_static")
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint64), planes[1].cpu().numpy().view(np.uint64))
    assert g.step_counts() == (1, 1)
    g.close()


# Wall sets of a 3-D box: every side (the reference's in.nml), the low sides,
# y and z low only (the two-level corner ghosts each time), and two without
# them.
WALL_SETS = {
    "all": (1, 1, 1, 1, 1, 1),
    "low": (1, 0, 1, 0, 1, 0),
    "ylzl": (0, 0, 1, 0, 1, 0),
    "x_yr_z": (1, 1, 0, 1, 1, 1),
    "xl_zl": (1, 0, 0, 0, 1, 0),
}


def _walled(walls):
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=22, ny=19, nz=16, dx=1.0, dy=1.0, dz=1.0, gammafix=5.0 / 3.0, cfl=0.5, endtime=1e30)
    (prm.no_outflow_xl, prm.no_outflow_xr, prm.no_outflow_yl, prm.no_outflow_yr, prm.no_outflow_zl,
     prm.no_outflow_zr) = (bool(w) for w in walls)
    return prm, corner_blob_ic(prm, seed=2)


def _corner_pairs(prm):
    """(corner ghost, its interior donor v) of the all-walls box: ghosts at
    y, z in {0, 1} and x in {0, 1, nx-2, nx-1}, donors at y, z in {3, 2} and
    x in {3, 2, nx-3, nx-4} (0-based)."""
    nx, ny = prm.nx, prm.ny

    def cid(x, y, z):
        return x + nx * (y + ny * z)

    return [(cid(xg, y, z), cid(xv, 3 - y, 3 - z)) for z in (0, 1) for y in (0, 1)
            for xg, xv in ((0, 3), (1, 2), (nx - 2, nx - 3), (nx - 1, nx - 4))]


def test_corner_ghost_lag_is_visible():
    """The oracle (the reference's step) on the all-walls corner-blob box:
    after 10 steps some corner ghost's energy differs from its donor's (it
    holds the donor's corrector result before source(1)), so a plain copy of
    the donor would not give the reference's state; their densities agree."""
    prm, ic = _walled(WALL_SETS["all"])
    og = po.OracleGrid(prm, ic)
    og.run(10)
    pairs = _corner_pairs(prm)
    assert any(og.cons[g, 4] != og.cons[v, 4] for g, v in pairs)
    assert all(og.cons[g, 0] == og.cons[v, 0] for g, v in pairs)


@pytest.mark.gpu
@pytest.mark.parametrize("walls", list(WALL_SETS))
def test_gpu_walled_corner_blob_matches_oracle(walls):
    """Fused steps on walled 3-D boxes whose fastest cells feed the corner
    ghosts: dt sequence and every cell's bits (ghosts included) equal the
    oracle's run of the reference step, with the fused kernels after the
    first step."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    prm, ic = _walled(WALL_SETS[walls])
    steps = 10
    og = po.OracleGrid(prm, ic)
    odts, _ = og.run(steps)
    print(f"case corner blob walls={walls}", flush=True)
    g = HydroGrid(prm, ic)
    dts = g.run(steps, mode=nv.BX_STEP_FUSED)
    assert list(dts) == list(odts)
    out = g.download()
    for k in ("cons", "u", "p"):
        assert np.array_equal(out[k].view(np.uint64), getattr(og, k).view(np.uint64)), k
    assert g.step_counts() == (steps - 1, 1) and g.fallback_count() == 0
    g.close()


@pytest.mark.gpu
def test_gpu_mixed_form_walled_box():
    """The mixed form on the all-walls 3-D box takes the fused kernels after
    the first step (its corner ghosts copy their donors) and stays within its
    error budget of the FP64 step: relative L1 of rho and p over the interior
    cells <= 1e-6 after 20 steps.  (The corner ghosts' pressure follows the
    exact sequence's stale velocity in the FP64 step only.)"""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.ics import boundary_flags

    from boxzyhydro_amd import ics

    prm, _ = _walled(WALL_SETS["all"])
    ic = ics.box_ic(prm, seed=2)
    steps = 20
    out = {}
    for prec in (nv.BX_PREC_FP64, nv.BX_PREC_MIXED):
        g = HydroGrid(prm, ic, precision=prec)
        g.run(steps, mode=nv.BX_STEP_FUSED)
        out[prec] = g.download()
        assert g.step_counts() == (steps - 1, 1) and g.fallback_count() == 0
        g.close()
    a, b = out[nv.BX_PREC_FP64], out[nv.BX_PREC_MIXED]
    m = boundary_flags(prm).reshape(-1) == 0
    for name, x, y in (("rho", a["cons"][m, 0], b["cons"][m, 0]), ("p", a["p"][m], b["p"][m])):
        l1 = np.abs(y - x).sum() / np.abs(x).sum()
        print(f"  {name}: relative L1 {l1:.2e}", flush=True)
        assert l1 <= 1e-6, name
