"""Parity at the benchmark's full sizes.

* 256^3 box, 100 steps (the 3-D box workload: outflow, gamma=5/3, smoothed
  seeded density/pressure, rms Mach 0.5) against the reference's own OpenMP
  step (tests/golden/box256.json, made by tests/golden/make_box256_golden.py
  with oracle/_ref): every dt and the final cons/u/p bit-identical, in both
  step modes.
* 512^3 box (137 M cells with ghosts, 64-bit plane offsets beyond 2^30
  elements): the fused step equals the exact call sequence bit for bit.
"""
import json
import os

import numpy as np
import pytest

from cases import case_inputs, sha

HERE = os.path.dirname(os.path.abspath(__file__))
BOX256 = os.path.join(HERE, "golden", "box256.json")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BOX256), reason="tests/golden/box256.json not generated")
def test_gpu_box256_100_steps_match_reference():
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = json.load(open(BOX256))["box256_100"]
    prm, ic = case_inputs(rec)
    for mode in (nv.BX_STEP_FUSED, nv.BX_STEP_EXACT):
        print(f"box 256^3 x {rec['steps']} steps, mode {mode}", flush=True)
        g = HydroGrid(prm, ic, device=0)
        dts = g.run(rec["steps"], mode=mode)
        assert [float(x) for x in dts] == rec["dt"]
        out = g.download()
        for k in ("cons", "u", "p"):
            assert sha(out[k]) == rec["final_sha256"][k], (mode, k)
        g.close()


def _smooth_box(prm):
    """Analytic smooth 3-D fields (no filtering, so that 137 M cells generate
    in seconds): density and pressure ripples and a swirling velocity."""
    from boxzyhydro_amd import ics

    x, y, z = ics.cell_centres(prm)
    k = 2.0 * np.pi / (prm.nx * prm.dx)
    sx, cx = np.sin(k * x)[None, None, :], np.cos(k * x)[None, None, :]
    sy, cy = np.sin(k * y)[None, :, None], np.cos(k * y)[None, :, None]
    sz, cz = np.sin(k * z)[:, None, None], np.cos(k * z)[:, None, None]
    rho = 1.0 + 0.1 * (sx * cy + 0.5 * cz)
    p = 1.0 + 0.05 * (cx * sy * sz + 0.3 * sx)
    vx = 0.3 * sy * cz
    vy = 0.3 * sz * cx + 0.1 * sx
    vz = 0.3 * sx * cy
    shp = prm.shape
    return ics._assemble(prm, np.broadcast_to(rho, shp), np.broadcast_to(vx, shp), np.broadcast_to(vy, shp),
                         np.broadcast_to(vz, shp), np.broadcast_to(p, shp))


@pytest.mark.gpu
def test_gpu_512_cube_fused_equals_exact():
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = box_params(512)
    print(f"box 512^3: {prm.ngrid} cells", flush=True)
    ic = _smooth_box(prm)
    ge = HydroGrid(prm, ic, device=0)
    de = ge.run(3, mode=nv.BX_STEP_EXACT)
    ce = ge.download()["cons"]
    ge.close()
    del ge
    gf = HydroGrid(prm, ic, device=0)
    del ic
    df = gf.run(3, mode=nv.BX_STEP_FUSED)
    cf = gf.download()["cons"]
    gf.close()
    assert list(df) == list(de)
    assert np.array_equal(cf.view(np.uint64), ce.view(np.uint64))
    assert np.isfinite(cf).all()
