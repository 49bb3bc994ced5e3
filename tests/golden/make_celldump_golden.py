"""Golden vectors of the celldump writer and restart reader, from the
reference's own write_files() and read_hydro() (compiled from its sources by
`make -C oracle ref`, driven by oracle/ref_harness.f90).

For each case: run the reference step `steps` times from the case's set-up,
let write_files(step) write the celldump (binary and ASCII), then restart
from the binary one with read_hydro(); record the sha256 and size of both
files and the sha256 of the restart's cons/u/p.  Writes celldump.json.
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import pyoracle as po  # noqa: E402
from cases import case_inputs, sha  # noqa: E402

CASES = {"box12_outflow": 5, "stress8": 3, "flatz_walls": 4, "sod_step1": 1, "sod_t02": 752}


def main():
    golden = json.load(open(os.path.join(HERE, "golden.json")))
    out = {}
    for name, steps in CASES.items():
        prm, ic = case_inputs(golden[name])
        b = po.run_reference(prm, ic, steps, threads=1, celldump=1)
        a = po.run_reference(prm, ic, steps, threads=1, celldump=2)
        rs = po.reference_restart(prm, b["celldump"], steps)
        out[name] = {"steps": steps, "time": float(b["time"][-1]),
                     "binary": {"sha256": hashlib.sha256(b["celldump"]).hexdigest(), "bytes": len(b["celldump"])},
                     "ascii": {"sha256": hashlib.sha256(a["celldump"]).hexdigest(), "bytes": len(a["celldump"])},
                     "restart_sha256": {k: sha(rs[k]) for k in ("cons", "u", "p")}}
        print(name, out[name]["binary"]["bytes"], out[name]["ascii"]["bytes"], flush=True)
    with open(os.path.join(HERE, "celldump.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
