"""Benchmark of the hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]

Workload: the 3-D box of 256^3 interior cells in FP64 (gamma=5/3, CFL 0.5,
outflow, seeded smoothed-Gaussian density/pressure, rms Mach 0.5 velocities),
stepped with the reference's full main-loop step (both RK stages, both
half-step sources, the Courant reduction).  A step is one cell-update of every
interior cell.  Inputs are resident in HBM before the timed region.

Prints one JSON line with the cell-updates/s, the HBM-roofline of the dominant
kernel measured with HIP events on the library's stream, and a CPU baseline:
the reference's own Fortran subroutines (oracle/_ref) timed on the host on a
bounded sample, or the C restatement of the oracle when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
B_ALG_FP64 = 200.0         # algorithmic bytes per cell-update (5 fields x 8 B x 5 streams)
B_STAGE2_FP64 = 120.0      # stage-2 kernel: read U*, read U^n, write U^{n+1}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(prm, ic, steps: int, threads: int):
    """Time the reference step on the host: the reference's own subroutines if
    built (oracle/_ref), else the C restatement (oracle/build)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    cells = prm.interior_cells * steps
    if po.have_reference():
        try:
            wd = "/dev/shm" if os.path.isdir("/dev/shm") else None
            r = po.run_reference(prm, ic, steps, threads=threads, workdir=wd)
            return {"value": cells / r["wall_s"], "unit": "cell-updates/s", "cores": r["threads"],
                    "kind": "reference",
                    "sample": f"{steps} steps of the same {prm.nx - 4}^3 box, the reference's Fortran "
                              f"subroutines (amdflang -O3 -fopenmp) driven through main.f90's step, "
                              f"wall time of the step loop"}
        except Exception as e:  # fall back to the restatement
            log(f"reference baseline unavailable: {e}")
    po.set_threads(threads)
    og = po.OracleGrid(prm, ic)
    t0 = time.perf_counter()
    og.run(steps)
    dt = time.perf_counter() - t0
    return {"value": cells / dt, "unit": "cell-updates/s", "cores": po.get_threads(), "kind": "port",
            "sample": f"{steps} steps of the same {prm.nx - 4}^3 box, C/OpenMP restatement of the reference "
                      f"step (oracle/boxzy_oracle.c)"}


def mixed_study(prm, ic, steps, warmup, sync):
    """The mixed form against the FP64 step on the same box: cell-updates/s of
    each and the per-field L1/Linf error of the mixed state after `steps`."""
    import numpy as np

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    res = {"workload": f"box {prm.nx - 4}^3, {steps} steps", "fields": {}}
    states = {}
    for name, prec in (("fp64", nv.BX_PREC_FP64), ("mixed", nv.BX_PREC_MIXED)):
        g = HydroGrid(prm, ic, device=0, precision=prec)
        g.run(1, mode=nv.BX_STEP_FUSED)  # first step from the set-up (exact)
        g.sync()
        t0 = time.perf_counter()
        g.run(steps - 1, mode=nv.BX_STEP_FUSED)
        g.sync()
        dt = time.perf_counter() - t0
        res[name] = {"cell_updates_per_s": prm.interior_cells * (steps - 1) / dt, "time": g.time,
                     "fallbacks": g.fallback_count()}
        states[name] = g.download()
        del g
    interior = ics.boundary_flags(prm).reshape(-1) == 0
    a, b = states["fp64"], states["mixed"]
    fields = {"rho": (a["cons"][:, 0], b["cons"][:, 0]), "p": (a["p"], b["p"]),
              "mx": (a["cons"][:, 1], b["cons"][:, 1]), "E": (a["cons"][:, 4], b["cons"][:, 4])}
    for k, (x, y) in fields.items():
        x, y = x[interior], y[interior]
        res["fields"][k] = {"l1_rel": float(np.abs(x - y).sum() / np.abs(x).sum()),
                            "linf_rel": float(np.abs(x - y).max() / np.abs(x).max())}
    res["speedup"] = res["mixed"]["cell_updates_per_s"] / res["fp64"]["cell_updates_per_s"]
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--mode", choices=["fused", "exact"], default="fused")
    ap.add_argument("--tile", type=int, nargs=3, default=None, metavar=("YCHUNK", "ZCHUNK", "OCC"))
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision line")
    ap.add_argument("--mixed-study", action="store_true",
                    help="mixed form vs FP64 at --size for --steps steps: speed and per-field error")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    if args.copy_calib:
        import torch

        n = args.copy_calib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8, "launches": 5}))
        return 0

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    try:
        import torch

        sync = torch.cuda.synchronize
    except Exception:  # torch is plumbing only
        def sync():
            return None

    prm = box_params(args.size)
    log(f"generating box {args.size}^3 ...")
    ic = ics.box_ic(prm, seed=0)
    if args.mixed_study:
        print(json.dumps(mixed_study(prm, ic, args.steps, args.warmup, sync)))
        return 0
    mode = nv.BX_STEP_FUSED if args.mode == "fused" else nv.BX_STEP_EXACT
    g = HydroGrid(prm, ic, device=0)
    if args.tile:
        g.tune(*args.tile)
    log(f"warmup {args.warmup} steps ({args.mode}) ...")
    g.run(args.warmup, mode=mode)
    g.sync()
    sync()
    g.profile(1)
    log(f"timing {args.steps} steps ...")
    t0 = time.perf_counter()
    dts = g.run(args.steps, mode=mode)
    g.sync()
    sync()
    t1 = time.perf_counter()
    s1_ms, s2_ms, nprof = g.profile(0)
    elapsed = t1 - t0
    if len(dts) != args.steps:
        raise RuntimeError(f"took {len(dts)} of {args.steps} steps")
    cells = prm.interior_cells
    value = cells * args.steps / elapsed
    out = {
        "metric": "cell-updates/s @256^3 FP64 vs mixed (HBM roofline %); adjoint/forward cost",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D box {args.size}^3 interior FP64, outflow, gamma=5/3, CFL 0.5, "
                               f"seeded smoothed Gaussian rho/p, rms Mach 0.5; full main-loop step",
                   "grid": [prm.nx, prm.ny, prm.nz], "mode": args.mode,
                   "fallback_steps": g.fallback_count()},
        "step_hbm_roofline_frac": value * B_ALG_FP64 / HBM_PEAK,
    }
    if nprof > 0:
        avg2 = s2_ms / nprof * 1e-3
        avg1 = s1_ms / nprof * 1e-3
        achieved = B_STAGE2_FP64 * cells / avg2
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_stage2_box{args.size}.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        out["roofline"] = {"kernel": "k_stage<2> (corrector stage)", "bound": "hbm", "achieved": achieved / 1e9,
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                           "traffic": traffic, "alg_bytes_per_launch": B_STAGE2_FP64 * cells,
                           "avg_launch_ms": avg2 * 1e3, "stage1_avg_launch_ms": avg1 * 1e3}
    else:
        out["roofline"] = None
    del g
    if not args.no_mixed:
        log(f"mixed form: {args.steps} steps ...")
        gm = HydroGrid(prm, ic, device=0, precision=nv.BX_PREC_MIXED)
        gm.run(args.warmup, mode=nv.BX_STEP_FUSED)
        gm.sync()
        gm.profile(1)
        t0 = time.perf_counter()
        gm.run(args.steps, mode=nv.BX_STEP_FUSED)
        gm.sync()
        tm = time.perf_counter() - t0
        m1, m2, mn = gm.profile(0)
        vm = cells * args.steps / tm
        out["mixed"] = {"value": vm, "unit": "cell-updates/s", "dtype": "f32 storage+reconstruction, f64 update",
                        "speedup_vs_fp64": vm / value, "ms_per_step": tm * 1e3 / args.steps,
                        "hbm_roofline_frac": vm * 100.0 / HBM_PEAK,
                        "stage_avg_launch_ms": [m1 / max(mn, 1), m2 / max(mn, 1)]}
        del gm
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_steps} steps, {args.cpu_threads} threads) ...")
        out["cpu_baseline"] = cpu_baseline(prm, ic, args.cpu_steps, args.cpu_threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
