"""The oracle (C restatement, oracle/boxzy_oracle.c) against the golden vectors
written by the reference's own Fortran subroutines (tests/golden/golden.json).

Every comparison is bitwise: dt values exactly, arrays through SHA-256 of their
float64 bytes."""
import numpy as np
import pytest

import pyoracle as po
from cases import Driver, OracleAdapter, case_inputs, ic_digest, sha

CASES = ["sod_step1", "box12_outflow", "box12_walls", "box24_outflow", "stress8", "flatz_walls",
         "riemann4_0", "riemann4_1", "riemann4_2", "riemann4_3", "sod_t02", "readme_sod_y"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(golden, name):
    rec = golden[name]
    prm, ic = case_inputs(rec)
    assert ic_digest(ic) == rec["ic_sha256"], "input generator changed"
    og = po.OracleGrid(prm, ic)
    if "snaps" in rec:
        got = {}

        def snap(tag, dt):
            got[tag] = (dt, {k: sha(v) for k, v in OracleAdapter(og).arrays().items()})

        got[0] = (None, {k: sha(v) for k, v in OracleAdapter(og).arrays().items()})
        dts, _ = Driver(OracleAdapter(og)).run(rec["steps"], prm.endtime, snap=snap)
        for s in rec["snaps"]:
            dt, digests = got[s["tag"]]
            if s["tag"] > 0:
                assert dt == s["dt"], f"snapshot {s['tag']}: dt"
            for k, v in s["sha256"].items():
                assert digests[k] == v, f"snapshot {s['tag']}: {k}"
    else:
        dts, _ = og.run(rec["steps"])
    assert len(dts) == rec["steps"]
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(getattr(og, k)) == rec["final_sha256"][k], k


def test_sod_profile_fixture(golden):
    """The stored x-profile of the reference's Sod run equals the oracle's."""
    rec = golden["sod_t02"]
    prm, ic = case_inputs(rec)
    og = po.OracleGrid(prm, ic)
    og.run(rec["steps"])
    row = og.cons.reshape(prm.nz, prm.ny, prm.nx, 5)[0, 2]
    assert np.array_equal(row[:, 0], np.array(rec["profile_iy3_iz1"]["rho"]))
    assert np.array_equal(og.p.reshape(prm.nz, prm.ny, prm.nx)[0, 2], np.array(rec["profile_iy3_iz1"]["p"]))


@pytest.mark.skipif(not po.have_reference(), reason="reference driver not built here")
def test_oracle_vs_reference_live():
    """When the reference build is present: a fresh seeded box, compared live."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd.params import box_params

    prm = box_params(10)
    prm.no_outflow_yl = prm.no_outflow_yr = True
    ic = ics.box_ic(prm, seed=11, sigma_cells=1.5)
    ref = po.run_reference(prm, ic, 6, threads=1)
    og = po.OracleGrid(prm, ic)
    dts, _ = og.run(6)
    assert list(dts) == list(ref["dt"])
    for k in ("cons", "u", "p"):
        assert np.array_equal(getattr(og, k).view(np.uint64), ref[k].view(np.uint64)), k
