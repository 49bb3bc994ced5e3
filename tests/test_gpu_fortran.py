"""The Fortran driver (boxzyhydro_amd/fortran: the reference's namelist and
main-loop call sequence over the C-ABI through ISO_C_BINDING) against the
reference's golden vectors."""
import os
import subprocess

import numpy as np
import pytest

from boxzyhydro_amd import ics
from cases import case_inputs, sha
from conftest import ROOT

DRIVER = os.path.join(ROOT, "boxzyhydro_amd", "fortran", "boxzy_driver")


def run_driver(tmp_path, prm, ic, mode, steps=1000000, restart=0):
    """Run the driver in tmp_path (its celldump files land there)."""
    nml = tmp_path / "run.nml"
    text = prm.to_namelist()
    if restart:
        text = text.replace("irestart=0", f"irestart={restart}")
    nml.write_text(text)
    icf = "-"
    if ic is not None:
        icf = str(tmp_path / "ic.bin")
        ics.write_ic_file(icf, ic)
    out = str(tmp_path / "out")
    subprocess.run([DRIVER, str(nml), icf, out, str(steps), mode], check=True, timeout=300, cwd=str(tmp_path))
    dt = np.fromfile(out + ".dt", np.float64).reshape(-1, 2)
    fin = np.fromfile(out + ".final", np.float64)
    n = prm.ngrid
    return dt[:, 0], {"cons": fin[:5 * n].reshape(n, 5), "u": fin[5 * n:8 * n].reshape(n, 3),
                      "p": fin[8 * n:9 * n]}


def test_driver_is_built():
    assert os.path.exists(DRIVER), "build with python -m boxzyhydro_amd.build"


@pytest.mark.gpu
@pytest.mark.parametrize("name,mode", [("sod_t02", "calls"), ("sod_t02", "fused"), ("box12_walls", "calls"),
                                       ("box24_outflow", "fused"), ("stress8", "calls")])
def test_driver_matches_reference(golden, tmp_path, name, mode):
    rec = golden[name]
    prm, ic = case_inputs(rec)
    print(f"fortran driver {name} {mode}", flush=True)
    dts, out = run_driver(tmp_path, prm, ic, mode, rec["steps"])
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k


@pytest.mark.gpu
def test_driver_builtin_setup_matches_reference_readme_run(golden, tmp_path):
    """No IC file: the driver's own init_conditions (Sod along y, frame cells)
    on the reference's shipped namelist values reproduces the reference run."""
    rec = golden["readme_sod_y"]
    prm, _ = case_inputs(rec)
    print("fortran driver builtin set-up", flush=True)
    dts, out = run_driver(tmp_path, prm, None, "calls")
    assert list(dts) == rec["dt"]
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["calls", "fused"])
def test_driver_writes_reference_celldumps_and_restarts(golden, tmp_path, mode):
    """write_files on main.f90:548-553's cadence (dtout > endtime: after step 1
    and at endtime) gives the reference's celldump files byte for byte, and a
    restart from the last one (read_hydro) gives the reference's restart state."""
    import hashlib
    import json

    cd = json.load(open(os.path.join(ROOT, "tests", "golden", "celldump.json")))
    rec = golden["sod_t02"]
    prm, ic = case_inputs(rec)
    print(f"fortran driver celldumps {mode}", flush=True)
    run_driver(tmp_path, prm, ic, mode, rec["steps"])
    files = sorted(f.name for f in tmp_path.glob("celldump.*"))
    assert files == ["celldump.00000001", f"celldump.{rec['steps']:08d}"]

    def digest(name):
        return hashlib.sha256((tmp_path / name).read_bytes()).hexdigest()

    assert digest("celldump.00000001") == cd["sod_step1"]["binary"]["sha256"]
    assert digest(files[-1]) == cd["sod_t02"]["binary"]["sha256"]
    # restart: no further step (maxsteps = irestart), the state read_hydro leaves
    _, out = run_driver(tmp_path, prm, None, mode, rec["steps"], restart=rec["steps"])
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == cd["sod_t02"]["restart_sha256"][k], k
