"""Tabulated H2 equation of state (H2STAT = 0, 1) against the reference's own
eos.f90 / state.f90 / main-loop step (golden vectors in tests/golden/eos.json,
made by tests/golden/make_eos_golden.py with oracle/_ref):

* the host table builder: every table bit-identical, small and default sizes;
* one state() call on the GPU (3-D and the nz=1 disk branch): cons, p,
  adindx, muc_array bit-identical;
* full steps on the GPU: every dt and the final state bit-identical, and the
  celldump file byte-identical (muc_array per cell).
"""
import dataclasses
import hashlib
import json
import os

import numpy as np
import pytest

from cases import case_inputs, sha

HERE = os.path.dirname(os.path.abspath(__file__))
EOS = json.load(open(os.path.join(HERE, "golden", "eos.json")))
TABLE_KEYS = ("rho_table", "deng_eos_array", "tk_table", "eng_table", "gamma_table", "muc_table",
              "eng_table2", "tk_table2", "muc_table2", "gamma_table2")


def _prm(rec):
    from boxzyhydro_amd.params import HydroParams

    return HydroParams(**rec["params"])


@pytest.mark.parametrize("name", sorted(EOS["tables"]))
def test_tables_match_reference(name):
    from boxzyhydro_amd.hydro import eos_tables

    rec = EOS["tables"][name]
    t = eos_tables(_prm(rec))
    assert t["NEOS_T"] == rec["NEOS_T"]
    assert t["log_rho_eos_low"] == rec["log_rho_eos_low"] and t["drho_eos"] == rec["drho_eos"]
    for k in TABLE_KEYS:
        assert sha(np.ascontiguousarray(t[k])) == rec["sha256"][k], k


@pytest.mark.parametrize("h2", [0, 1])
def test_tables_match_reference_other_mixtures(h2):
    """Abundances with metals, another ortho:para ratio and other molecular
    constants, against oracle/_ref directly (build container only)."""
    import pyoracle as po

    from boxzyhydro_amd.hydro import eos_tables
    from boxzyhydro_amd.params import box_params

    if not po.have_reference():
        pytest.skip("oracle/_ref not built")
    prm = dataclasses.replace(box_params(8), H2STAT=h2, NEOS_RHO=7, dTk_eos=37.0, tk_bgrnd=3.0,
                              xabun=0.70, yabun=0.28, zabun=0.02, mu_z=14.0, ac=1.0, bc=1.0,
                              brot=80.0, vib=6000.0)
    ref = po.reference_eos_tables(prm)
    ours = eos_tables(prm)
    for k in TABLE_KEYS:
        assert np.array_equal(ours[k], ref[k]), k


def test_tables_physics_limits():
    """What the physics fixes, independent of the reference: energy rises with
    T in every density row; muc lies between the fully ionized and the fully
    molecular limits; H2 + He with rotation frozen (T <= 15 K) has gamma near
    5/3, and with rotation excited and vibration frozen (400-600 K) the
    closed form 1 + (x/2 + y/4) / (5/2 x/2 + 3/2 y/4); the log-energy grid of
    table 2 starts at the first energy of table 1."""
    from boxzyhydro_amd.hydro import eos_tables
    from boxzyhydro_amd.params import box_params

    prm = dataclasses.replace(box_params(8), H2STAT=1, NEOS_RHO=12, dTk_eos=5.0, tk_bgrnd=10.0)
    t = eos_tables(prm)
    assert np.all(np.diff(t["eng_table"], axis=1) > 0)
    x, y, z, mz = prm.xabun, prm.yabun, prm.zabun, prm.mu_z
    mu_ion = 1.0 / (x * 2.0 + y * 0.25 + z / mz)
    mu_mol = 1.0 / (x * 0.5 + y * 0.25 + z / mz)
    assert np.all(t["muc_table"] >= mu_ion * (1 - 1e-12)) and np.all(t["muc_table"] <= mu_mol * (1 + 1e-12))
    tk = t["tk_table"]
    cold = tk <= 15
    warm = (tk >= 400) & (tk <= 600)
    g = t["gamma_table"]
    assert np.all(np.abs(g[:, cold] - 5 / 3) < 0.015)
    g_rot = 1.0 + (x / 2 + y / 4) / (2.5 * x / 2 + 1.5 * y / 4)
    assert np.all(np.abs(g[:, warm] - g_rot) < 0.003)
    assert np.all(t["eng_table2"][:, 0] >= t["eng_table"][:, 0])


def test_tables_reject_bad_sizes():
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import eos_tables
    from boxzyhydro_amd.params import box_params

    with pytest.raises(nv.BoxzyError, match="NEOS_RHO"):
        eos_tables(dataclasses.replace(box_params(8), H2STAT=0, NEOS_RHO=1))
    with pytest.raises(nv.BoxzyError, match="dTk_eos"):
        eos_tables(dataclasses.replace(box_params(8), H2STAT=0, dTk_eos=0.0))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EOS["state"]))
def test_gpu_state_matches_reference(name):
    from boxzyhydro_amd.hydro import HydroGrid

    rec = EOS["state"][name]
    prm, ic = case_inputs(rec)
    print(f"eos state {name}", flush=True)
    g = HydroGrid(prm, ic, device=0)
    g.state()
    out = g.download()
    out["muc"] = g.download_muc()
    g.close()
    for k in ("cons", "p", "adindx", "muc"):
        assert sha(out[k]) == rec["sha256"][k], k


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "fused"])
@pytest.mark.parametrize("name", sorted(EOS["runs"]))
def test_gpu_steps_match_reference(name, mode, tmp_path):
    """Full steps with the tables; the fused mode runs the exact sequence
    with a tabulated EOS (its kernels assume a fixed gamma)."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = EOS["runs"][name]
    prm, ic = case_inputs(rec)
    print(f"eos run {name} {mode}", flush=True)
    g = HydroGrid(prm, ic, device=0)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_EXACT if mode == "exact" else nv.BX_STEP_FUSED)
    assert [float(x) for x in dts] == rec["dt"]
    out = g.download()
    out["muc"] = g.download_muc()
    for k in ("cons", "u", "p", "adindx", "muc"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    path = str(tmp_path / "celldump")
    g.write_celldump(path, rec["steps"])
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == rec["celldump_sha256"]
    g.close()


@pytest.mark.gpu
def test_gpu_fortran_driver_tabulated_eos(tmp_path):
    """The Fortran driver reads H2STAT and the table keys from &eos_input and
    calls bx_eos_init (main.f90:70-71); call-by-call steps match."""
    from test_gpu_fortran import run_driver

    rec = EOS["runs"]["cloud_walls_h2stat0"]
    prm, ic = case_inputs(rec)
    print("fortran driver tabulated eos", flush=True)
    dts, out = run_driver(tmp_path, prm, ic, "calls", rec["steps"])
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k


@pytest.mark.gpu
def test_gpu_clearing_gravity_keeps_tabulated_eos_on_exact_steps():
    """Switching gravity off on a tabulated-EOS context must not re-enable the
    fixed-gamma fused kernels: the run still matches the reference."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = EOS["runs"]["cloud16_h2stat0"]
    prm, ic = case_inputs(rec)
    g = HydroGrid(prm, ic, device=0)
    g.set_gravity(nv.BX_GRAV_POTENTIAL, None, initial=False)
    g.set_gravity(0)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_FUSED)
    assert [float(x) for x in dts] == rec["dt"]
    assert sha(g.download()["cons"]) == rec["final_sha256"]["cons"]
    g.close()
