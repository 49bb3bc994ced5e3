"""The angular-momentum flux (fluxangmom=.true.: flux_ang.f90:6-618 in place
of flux.f90, and the source.f90:73-90 terms) against the reference's own
code (golden vectors in tests/golden/ang.json, made by
tests/golden/make_ang_golden.py with oracle/_ref): every dt and the final
state bit-identical, and the state after every sub-call of step 1, cons_old
included (flux_ang transforms cons_old and back when avg=1)."""
import json
import os

import pytest

from cases import Driver, HipAdapter, case_inputs, sha

HERE = os.path.dirname(os.path.abspath(__file__))
ANG = json.load(open(os.path.join(HERE, "golden", "ang.json")))


def test_namelist_carries_fluxangmom_and_yoffset():
    """The reference reads fluxangmom and yoffset from &grid_params (input.f90:8)."""
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(**ANG["ang_box_yoffset"]["params"])
    text = prm.to_namelist()
    assert "fluxangmom=.true." in text and "yoffset=0.7d0" in text
    back = HydroParams.from_namelist(text)
    assert back.fluxangmom is True and back.yoffset == 0.7


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(ANG))
def test_gpu_steps_match_reference(name):
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = ANG[name]
    prm, ic = case_inputs(rec)
    print(f"flux_ang case {name} grid {prm.nx}x{prm.ny}x{prm.nz}", flush=True)
    g = HydroGrid(prm, ic, device=0)
    # fused mode runs the exact sequence with fluxangmom (the fused kernels flux linear momentum)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_FUSED)
    assert [float(x) for x in dts] == rec["dt"]
    out = g.download()
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(n for n in ANG if "snaps" in ANG[n]))
def test_gpu_subroutines_match_reference_snapshots(name):
    from boxzyhydro_amd.hydro import HydroGrid

    rec = ANG[name]
    prm, ic = case_inputs(rec)
    print(f"flux_ang snapshots {name}", flush=True)
    ad = HipAdapter(HydroGrid(prm, ic, device=0))
    got = {0: (None, {k: sha(v) for k, v in ad.arrays().items()})}

    def snap(tag, dt):
        got[tag] = (dt, {k: sha(v) for k, v in ad.arrays().items()})

    dts, _ = Driver(ad).run(rec["steps"], prm.endtime, snap=snap)
    for s in rec["snaps"]:
        dt, digests = got[s["tag"]]
        if s["tag"] > 0:
            assert dt == s["dt"], f"snapshot {s['tag']}: dt"
        for k, v in s["sha256"].items():
            assert digests[k] == v, f"snapshot {s['tag']}: {k}"
    assert list(dts) == rec["dt"]


@pytest.mark.gpu
def test_gpu_fortran_driver_fluxangmom(tmp_path):
    """The Fortran driver passes fluxangmom and yoffset from &grid_params."""
    from test_gpu_fortran import run_driver

    rec = ANG["ang_box_yoffset"]
    prm, ic = case_inputs(rec)
    print("fortran driver flux_ang", flush=True)
    dts, out = run_driver(tmp_path, prm, ic, "calls", rec["steps"])
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
