"""Write the golden vectors of the angular-momentum flux (fluxangmom=.true.:
flux_ang.f90 and the source.f90:73-90 terms) from the reference's own code
(oracle/_ref/ref_harness, which calls flux_ang as main.f90:343-369 does):
every dt, the final cons/u/p, and the state after every sub-call of step 1
(snapshots, cons_old included).

Only run in the build container (the reference is not available elsewhere):

    python tests/golden/make_ang_golden.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from boxzyhydro_amd.params import HydroParams  # noqa: E402
from cases import case_inputs  # noqa: E402
from make_golden import case_record  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ang.json")
OUTFLOW = dict(no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False, no_outflow_yr=False,
               no_outflow_zl=False, no_outflow_zr=False)


def record(prm, gen, steps, threads, mode):
    rec = {"params": {k: getattr(prm, k) for k in prm.__dataclass_fields__}, "generator": gen}
    _, ic = case_inputs(rec)
    out, _ = case_record(prm, ic, steps, threads, mode, gen)
    return out


def main() -> None:
    import pyoracle as po

    if not po.have_reference():
        raise SystemExit("build the reference driver first: make -C oracle ref")
    cases = {}
    # single-layer disk (the reference's own runs are nz=1), outflow
    disk = HydroParams(nx=28, ny=28, nz=1, dx=0.1, dy=0.1, dz=0.1, endtime=1e30, cfl=0.4, gammafix=1.4,
                       fluxangmom=True, **OUTFLOW)
    cases["ang_disk"] = record(disk, {"kind": "rotating_box", "seed": 2, "kw": {"omega": 1.0}}, 12, 4, 1)
    # 3-D box, z walls, axis shifted by yoffset (the transform uses y+yoffset,
    # the pressure projection of get_fluxes_ang uses y alone)
    box = HydroParams(nx=16, ny=18, nz=12, dx=0.2, dy=0.2, dz=0.25, endtime=1e30, cfl=0.4, gammafix=5.0 / 3.0,
                      fluxangmom=True, yoffset=0.7, **{**OUTFLOW, "no_outflow_zl": True, "no_outflow_zr": True})
    cases["ang_box_yoffset"] = record(box, {"kind": "rotating_box", "seed": 3,
                                            "kw": {"omega": 0.5, "yoffset": 0.7, "sigma_cells": 1.0}}, 10, 1, 1)
    # reflecting x/y walls, flat z
    walls = HydroParams(nx=20, ny=16, nz=4, dx=0.1, dy=0.1, dz=0.1, endtime=1e30, cfl=0.4, gammafix=1.4,
                        fluxangmom=True)
    cases["ang_walls_flatz"] = record(walls, {"kind": "rotating_box", "seed": 4, "kw": {"omega": 2.0}}, 8, 1, 1)
    # with the tabulated H2 EOS
    h2 = dataclasses.replace(disk, H2STAT=0, NEOS_RHO=60, dTk_eos=5.0, tk_bgrnd=10.0, dx=1.0, dy=1.0, dz=1.0,
                             cfl=0.5)
    cases["ang_h2_disk"] = record(h2, {"kind": "h2_disk", "seed": 6, "kw": {"omega": 0.05}}, 10, 4, 0)
    with open(OUT, "w") as f:
        json.dump(cases, f, indent=1)
    for k, r in cases.items():
        print(k, r["steps"], r["dt"][:2])
    print("wrote", OUT)


if __name__ == "__main__":
    main()
