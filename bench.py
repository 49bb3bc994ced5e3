"""Benchmark of the hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]

Workload: the 3-D box of 256^3 interior cells in FP64 (gamma=5/3, CFL 0.5,
outflow, seeded smoothed-Gaussian density/pressure, rms Mach 0.5 velocities),
stepped with the reference's full main-loop step (both RK stages, both
half-step sources, the Courant reduction).  A step is one cell-update of every
interior cell.  Inputs are resident in HBM before the timed region.

Prints one JSON line: the FP64 cell-updates/s (value), the HBM roofline of
the dominant kernel (the corrector's x/z kernel) timed with HIP events on the
library's stream, the mixed-precision form on the same box, the batched
shock tubes (512 problems, one persistent launch replayed from a hipGraph),
the tabulated H2 EOS (state() kernel rate, step cost against a fixed gamma),
flux_ang and gravity from a given potential (step costs),
the adjoint/forward cost (128^3 box, 200 steps, checkpoints every 20) and a
CPU baseline: the reference's own Fortran subroutines (oracle/_ref) timed on
the host on a bounded sample, or the C restatement when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12    # MI355X HBM3E, bytes/s
B_STEP_FP64 = 200.0  # algorithmic bytes per cell-update: U^n in, U* out, U*, U^n in, U^{n+1} out (5 x 8 B each)
B_STEP_MIXED = 100.0
B_XZ2_FP64 = 120.0   # corrector x/z kernel: read U*, read U^n, write U^{n+1} (5 FP64 fields each)
B_XZ2_MIXED = 60.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Physical cores in this process's affinity mask (one per core id and
    package), and the CPU model from lscpu."""
    import subprocess

    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            with open(f"{base}/core_id") as f:
                core = f.read().strip()
            with open(f"{base}/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    info = {"affinity_cpus": len(cpus), "physical_cores": len(cores)}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "NUMA node(s)"):
                info[k.strip()] = v.strip()
    except Exception as e:  # lscpu is informational only
        info["lscpu_error"] = str(e)
    return info


def cpu_baseline(prm, ic, steps: int, threads: int = 0):
    """Time the reference step on the host: the reference's own subroutines if
    built (oracle/_ref), else the C restatement (oracle/build).  Threads: the
    physical cores of the affinity mask (unless given), bound with
    OMP_PROC_BIND=close and OMP_PLACES=cores; two runs of `steps` steps, the
    hot path alone and with the per-step conservation diagnostics that the
    reference's main loop runs every step (main.f90:222-223)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    host = host_cores()
    threads = threads or host["physical_cores"]
    bind = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
    cells = prm.interior_cells * steps
    if po.have_reference():
        try:
            wd = "/dev/shm" if os.path.isdir("/dev/shm") else None
            r = po.run_reference(prm, ic, steps, threads=threads, workdir=wd, env_extra=bind)
            rd = po.run_reference(prm, ic, steps, threads=threads, workdir=wd, env_extra=bind, diag=1)
            return {"value": cells / r["wall_s"], "unit": "cell-updates/s", "cores": r["threads"],
                    "kind": "reference",
                    "with_diagnostics": cells / rd["wall_s"],
                    "sample": f"{steps} steps of the same {prm.nx - 4}^3 box, the reference's Fortran "
                              f"subroutines (amdflang -O3 -fopenmp) driven through main.f90's step, wall "
                              f"time of the step loop (omp_get_wtime); with_diagnostics adds "
                              f"conservation_diagnostic and get_etot every step",
                    "omp": dict(bind, OMP_NUM_THREADS=str(threads)), "host": host}
        except Exception as e:  # fall back to the restatement
            log(f"reference baseline unavailable: {e}")
    os.environ.update(bind)
    po.set_threads(threads)
    og = po.OracleGrid(prm, ic)
    t0 = time.perf_counter()
    og.run(steps)
    dt = time.perf_counter() - t0
    return {"value": cells / dt, "unit": "cell-updates/s", "cores": po.get_threads(), "kind": "port",
            "sample": f"{steps} steps of the same {prm.nx - 4}^3 box, C/OpenMP restatement of the reference "
                      f"step (oracle/boxzy_oracle.c)", "omp": dict(bind, OMP_NUM_THREADS=str(threads)),
            "host": host}


def timed_run(prm, ic, precision, steps, warmup, tile=None, mode=None):
    """(cell-updates/s, ms per step, per-kernel average ms [y1, xz1, y2, xz2], launches timed, fallbacks).

    The timed steps run in `mode` (default: graph replays); the per-kernel
    averages come from `steps` further fused steps timed with HIP events."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    mode = nv.BX_STEP_GRAPH if mode is None else mode
    g = HydroGrid(prm, ic, device=0, precision=precision)
    if tile:
        g.tune(*tile)
    g.run(warmup, mode=mode, endtime=1e300)
    g.sync()
    t0 = time.perf_counter()
    dts = g.run(steps, mode=mode, endtime=1e300)
    g.sync()
    elapsed = time.perf_counter() - t0
    if len(dts) != steps:
        raise RuntimeError(f"took {len(dts)} of {steps} steps")
    g.profile(1)
    g.run(steps, mode=nv.BX_STEP_FUSED, endtime=1e300)
    g.sync()
    ms, nprof = g.profile(0)
    per = [m / max(nprof, 1) for m in ms]
    fb = g.fallback_count()
    del g
    return prm.interior_cells * steps / elapsed, elapsed * 1e3 / steps, per, nprof, fb


def graph_cost(steps_small: int, steps_box: int, n: int):
    """Wall time per step of the fused step taken one at a time (a host
    synchronisation per step) and replayed from captured graphs, on the Sod
    tube (404x8x4 with ghosts, launch-bound) and on the n^3 box."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import HydroParams, box_params

    sod = HydroParams(nx=404, ny=8, nz=4, dx=0.0025, dy=0.0025, dz=0.0025, gammafix=1.4, cfl=0.5,
                      small_rho=1e-12, small_eps=1e-40, vlimit=10.0, endtime=1e300,
                      no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False, no_outflow_yr=False,
                      no_outflow_zl=False, no_outflow_zr=False)
    out = {}
    for name, prm, ic, steps in (("sod_404x8x4", sod, ics.sod_ic(sod, axis=0, x0=0.0), steps_small),
                                 (f"box_{n}^3", box_params(n), None, steps_box)):
        if ic is None:
            ic = ics.box_ic(prm, seed=0)
        row = {}
        for label, mode in (("one_at_a_time", nv.BX_STEP_FUSED), ("graph", nv.BX_STEP_GRAPH)):
            g = HydroGrid(prm, ic, device=0)
            g.run(3, mode=mode)
            g.sync()
            t0 = time.perf_counter()
            d = g.run(steps, mode=mode, record_dt=False)
            g.sync()
            el = time.perf_counter() - t0
            row[label + "_ms_per_step"] = el * 1e3 / steps
            row[label + "_cell_updates_per_s"] = prm.interior_cells * steps / el
            if mode == nv.BX_STEP_GRAPH:
                row["replayed_steps"] = g.graph_step_count()
            g.close()
        row["speedup"] = row["one_at_a_time_ms_per_step"] / row["graph_ms_per_step"]
        out[name] = row
    return out


def mixed_study(prm, ic, steps):
    """The mixed form against the FP64 step on the same box: cell-updates/s of
    each (graph replays after the first, exact step) and the per-field L1 and
    Linf error of the mixed state against the FP64 GPU run after `steps`
    steps, relative to the FP64 field's L1 norm and max magnitude."""
    import numpy as np

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    res = {"workload": f"box {prm.nx - 4}^3 (same generator), {steps} steps, FP64 and mixed from the same "
                       f"set-up", "fields": {}}
    states = {}
    for name, prec in (("fp64", nv.BX_PREC_FP64), ("mixed", nv.BX_PREC_MIXED)):
        g = HydroGrid(prm, ic, device=0, precision=prec)
        g.run(1, mode=nv.BX_STEP_FUSED)  # first step from the set-up (exact)
        g.sync()
        t0 = time.perf_counter()
        g.run(steps - 1, mode=nv.BX_STEP_GRAPH)
        g.sync()
        dt = time.perf_counter() - t0
        res[name] = {"cell_updates_per_s": prm.interior_cells * (steps - 1) / dt, "time": g.time,
                     "fallbacks": g.fallback_count()}
        states[name] = g.download()
        g.close()
        del g
    interior = ics.boundary_flags(prm).reshape(-1) == 0
    a, b = states["fp64"], states["mixed"]
    names = ("rho", "mx", "my", "mz", "E")
    for j, k in enumerate(names):
        x, y = a["cons"][interior, j], b["cons"][interior, j]
        d = np.abs(x - y)
        res["fields"][k] = {"l1_rel": float(d.sum() / np.abs(x).sum()), "linf_rel": float(d.max() / np.abs(x).max())}
    x, y = a["p"][interior], b["p"][interior]
    d = np.abs(x - y)
    res["fields"]["p"] = {"l1_rel": float(d.sum() / np.abs(x).sum()), "linf_rel": float(d.max() / np.abs(x).max())}
    res["speedup"] = res["mixed"]["cell_updates_per_s"] / res["fp64"]["cell_updates_per_s"]
    res["mixed"]["step_hbm_roofline_frac"] = res["mixed"]["cell_updates_per_s"] * B_STEP_MIXED / HBM_PEAK
    res["fp64"]["step_hbm_roofline_frac"] = res["fp64"]["cell_updates_per_s"] * B_STEP_FP64 / HBM_PEAK
    return res


def batched(n_tubes: int, reps: int):
    """512 shock tubes (L log-uniform in [64,4096], t_end=0.1) in one
    persistent launch replayed from a captured hipGraph, work split picked by
    the autotuner."""
    import numpy as np

    from boxzyhydro_amd import ics
    from boxzyhydro_amd.autotune import tune_batch, tune_gangs
    from boxzyhydro_amd.batch import TubeBatch

    tubes = ics.riemann_tubes(n_tubes, seed=0)
    b = TubeBatch(tubes)
    best, res = tune_batch(b, reps=1)
    single = b.result()  # one workgroup per tube: the bits a gang must reproduce
    gbest, gres = tune_gangs(b, reps=1)
    b.run(graph=True)
    b.sync()
    ms = min([(b.run(graph=True), b.sync())[1] for _ in range(reps)])
    out = b.result()
    same = all(np.array_equal(x[k], y[k]) for x, y in zip(out, single) for k in ("cons", "u", "p")) and \
        all(x["time"] == y["time"] and x["steps"] == y["steps"] and x["status"] == y["status"]
            for x, y in zip(out, single))
    steps = np.array([r["steps"] for r in out])
    row_cells = np.asarray(b.nx - 4)
    row_updates = float(np.sum(steps * row_cells))
    _, interior3 = b.cells()
    full_updates = float(np.sum(steps * row_cells * 16))  # each tube grid is L x 4 x 4 interior cells
    return {"workload": f"{n_tubes} shock tubes, L log-uniform in [64,4096] (x4x4 interior), gamma 1.4, "
                        f"t_end 0.1, FP64, one persistent launch replayed from a hipGraph, the costliest "
                        f"tubes each over a gang of workgroups",
            "ms": ms, "steps_min": int(steps.min()), "steps_max": int(steps.max()),
            "row_cell_updates_per_s": row_updates / (ms * 1e-3),
            "grid_equivalent_cell_updates_per_s": full_updates / (ms * 1e-3),
            "grid_equivalent_note": "the L x 4 x 4 grids' cell-updates; their 16 rows are identical and "
                                    "computed once, so this is not a throughput (row_cell_updates_per_s is)",
            "tile": {"waves_per_block": best[0], "blocks_per_cu": best[1]},
            "tile_ms": {f"{c[0]}x{c[1]}": t * 1e3 for c, t in res},
            "gangs": {"members": gbest[0], "tubes": gbest[1], "fallbacks": b.gang_fallbacks(),
                      "same_bits_as_one_workgroup_per_tube": bool(same),
                      "candidates_ms": {f"{c[0]}x{c[1]}": t * 1e3 for c, t in gres}},
            "longest_tube_steps": int(steps.max()),
            "status_bits": int(np.bitwise_or.reduce([r["status"] for r in out]))}


def box_workload(n_boxes: int, seed: int = 0, walls: bool = False):
    """Uneven 3-D boxes: interior cells per axis log-uniform in [8, 32] (each
    axis drawn separately), outflow (or, with ``walls``, walls on a random
    subset of sides), seeded smoothed rho/p with rms Mach 0.5, gamma 5/3,
    CFL 0.5, run to t_end = 2."""
    import numpy as np

    from boxzyhydro_amd import ics
    from boxzyhydro_amd.params import HydroParams

    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_boxes):
        nx, ny, nz = (int(round(v)) + 4 for v in np.exp(rng.uniform(np.log(8), np.log(32), 3)))
        wl = [bool(w) and walls for w in rng.integers(0, 2, 6)]
        prm = HydroParams(nx=nx, ny=ny, nz=nz, dx=0.25, dy=0.25, dz=0.25, gammafix=5.0 / 3.0, cfl=0.5,
                          endtime=2.0, no_outflow_xl=wl[0], no_outflow_xr=wl[1], no_outflow_yl=wl[2],
                          no_outflow_yr=wl[3], no_outflow_zl=wl[4], no_outflow_zr=wl[5])
        out.append((prm, ics.box_ic(prm, seed=seed * 100000 + i, sigma_cells=2.0)))
    return out


def boxes_cost(n_boxes: int, reps: int, one_at_a_time: bool = True, walls: bool = False):
    """The uneven-box workload in one persistent launch replayed from a
    captured hipGraph (work split picked by the autotuner), against the same
    boxes run one after another with bx_step's fused mode (the same bits; the
    exact sequence where a grid needs it)."""
    import numpy as np

    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.autotune import tune_boxes
    from boxzyhydro_amd.boxes import GridBatch
    from boxzyhydro_amd.hydro import HydroGrid

    grids = box_workload(n_boxes, walls=walls)
    b = GridBatch(grids)
    best, res = tune_boxes(b, reps=1)
    b.run(maxsteps=100000, graph=True)
    b.sync()
    ms = min([(b.run(maxsteps=100000, graph=True), b.sync())[1] for _ in range(reps)])
    out = b.result()
    steps = np.array([r["steps"] for r in out])
    interior = np.array([prm.interior_cells for prm, _ in grids])
    updates = float(np.sum(steps * interior))
    b.close()
    seq_s = 0.0
    for prm, ic in (grids if one_at_a_time else []):
        g = HydroGrid(prm, ic, device=0)
        g.sync()
        t0 = time.perf_counter()
        g.run(100000, mode=nv.BX_STEP_FUSED, record_dt=False)
        g.sync()
        seq_s += time.perf_counter() - t0
        g.close()
    return {"workload": f"{n_boxes} uneven 3-D boxes, interior per axis log-uniform in [8,32], "
                        f"{'random walls' if walls else 'outflow'}, gamma 5/3, CFL 0.5, t_end 2, FP64, "
                        f"one persistent launch replayed from a hipGraph",
            "ms": ms, "cells_interior": int(interior.sum()), "steps_min": int(steps.min()),
            "steps_max": int(steps.max()), "cell_updates_per_s": updates / (ms * 1e-3),
            "one_grid_at_a_time_s": seq_s if one_at_a_time else None,
            "one_grid_at_a_time_cell_updates_per_s": updates / seq_s if one_at_a_time else None,
            "speedup_vs_one_at_a_time": seq_s / (ms * 1e-3) if one_at_a_time else None,
            "tile": {"waves_per_block": best[0], "blocks_per_cu": best[1]},
            "tile_ms": {f"{c[0]}x{c[1]}": t * 1e3 for c, t in res},
            "status_bits": int(np.bitwise_or.reduce([r["status"] for r in out]))}


def adjoint_cost(n: int, nsteps: int, every: int):
    """Backward (recompute from checkpoints + reverse sweep) against the plain
    forward of the same frozen-dt map, on the adjoint workload."""
    import numpy as np
    import torch

    from boxzyhydro_amd import ics
    from boxzyhydro_amd.adjoint import AdjointStepper
    from boxzyhydro_amd.params import box_params

    prm = box_params(n)
    ic = ics.box_ic(prm, seed=0)
    tic = ics.box_ic(prm, seed=1)
    dev = torch.device("cuda", 0)
    st = AdjointStepper(prm, ic["u"], nsteps, every)
    rho0 = torch.as_tensor(ic["cons"][:, 0].copy(), device=dev)
    p0 = torch.as_tensor(ic["p"].copy(), device=dev)
    tst = AdjointStepper(prm, tic["u"], nsteps, 0)
    tgt = tst.forward(torch.as_tensor(tic["cons"][:, 0].copy(), device=dev),
                      torch.as_tensor(tic["p"].copy(), device=dev), checkpoint=False)
    del tst
    interior = torch.as_tensor(ics.boundary_flags(prm).reshape(-1) == 0, device=dev)
    t_courant = float("inf")
    for _ in range(2):  # Courant dt of every step (formed on the device), kept as the frozen sequence;
        t0 = time.perf_counter()  # the first call also allocates the checkpoints and segment stores
        st.forward(rho0, p0)
        t_courant = min(t_courant, time.perf_counter() - t0)
    dts = st.dts.copy()
    t_fwd = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        st.forward(rho0, p0, dts=dts, checkpoint=False)
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    res = {}
    for label, nbytes in (("kept", -1), ("recompute", 0)):
        st.store(nbytes)
        r0 = st.store(nbytes)
        t_fwd_ck = t_bwd = float("inf")
        for _ in range(2):  # forward with checkpoints, then its backward
            t0 = time.perf_counter()
            rT = st.forward(rho0, p0, dts=dts, checkpoint=True)
            t_fwd_ck = min(t_fwd_ck, time.perf_counter() - t0)
            lam = (rT - tgt) * interior
            t0 = time.perf_counter()
            g_rho, g_p = st.backward(lam)
            t_bwd = min(t_bwd, time.perf_counter() - t0)
        res[label] = (t_fwd_ck, t_bwd, (st.store(nbytes) - r0) // 2, g_rho, g_p)
    (t_fwd_ck, t_bwd, rec, g_rho, g_p), (t_fwd_ck0, t_bwd0, rec0, g_rho0, g_p0) = res["kept"], res["recompute"]
    return {"workload": f"box {n}^3 FP64, {nsteps} steps, checkpoints every {every}, frozen dt; "
                        f"loss 0.5|rho_T - rho_target|^2, gradient w.r.t. initial rho and p",
            "forward_ms": t_fwd * 1e3, "forward_courant_ms": t_courant * 1e3,
            "forward_checkpointed_ms": t_fwd_ck * 1e3, "backward_ms": t_bwd * 1e3,
            "adjoint_over_forward": t_bwd / t_fwd,
            "adjoint_over_forward_courant": t_bwd / t_courant,
            "segments_recomputed": rec,
            "recompute": {"forward_checkpointed_ms": t_fwd_ck0 * 1e3, "backward_ms": t_bwd0 * 1e3,
                          "adjoint_over_forward": t_bwd0 / t_fwd, "segments_recomputed": rec0,
                          "same_gradient_bits": bool(torch.equal(g_rho, g_rho0) and torch.equal(g_p, g_p0))},
            "note": "checkpoints every 20 steps; the forward also keeps the per-step states of as many "
                    "segments as half the free HBM holds (segments_recomputed says how many it could not), "
                    "'recompute' keeps one segment and recomputes the others from their checkpoints; "
                    "forward_ms is the same frozen-dt map without checkpoints (no per-step read-back), "
                    "forward_courant_ms the forward that forms each step's Courant dt on the device (with checkpoints)",
            "grad_norm": float(torch.sqrt((g_rho * g_rho).sum() + (g_p * g_p).sum())),
            "finite": bool(np.isfinite(float(g_rho.abs().max())))}


B_STATE_TAB = 64.0  # tabulated state(): read rho, E, u (5 x 8 B), write p, adindx, muc_array (3 x 8 B)


def walls_cost(n: int, steps: int, warmup: int, outflow_ms: float):
    """The headline box with walls on every side, the reference's shipped
    in.nml (utils.f90:338-403): the same graph-replayed fused step, whose
    ghost fill also writes the 16 two-level corner ghosts; ms per step
    against the outflow box's."""
    import dataclasses

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = dataclasses.replace(box_params(n), no_outflow_xl=True, no_outflow_xr=True, no_outflow_yl=True,
                              no_outflow_yr=True, no_outflow_zl=True, no_outflow_zr=True)
    ic = ics.box_ic(prm, seed=0)
    value, ms, _, _, fb = timed_run(prm, ic, nv.BX_PREC_FP64, steps, warmup)
    g = HydroGrid(prm, ic, device=0)
    g.run(3, mode=nv.BX_STEP_FUSED, endtime=1e300)
    fused, exact = g.step_counts()
    del g
    return {"workload": f"3-D box {n}^3 interior FP64, walls on all six sides, otherwise the headline box",
            "cell_updates_per_s": value, "ms_per_step": ms, "over_outflow": ms / outflow_ms,
            "fallback_steps": fb, "path_of_3_fused_mode_steps": {"fused": fused, "exact": exact}}


def host_cost(prm, ic, steps: int):
    """The step with the boundary's host hand-over included (PCIe): upload of
    the reference's host arrays (cons, u, p, adindx), `steps` graph-mode
    steps, download of cons, u and p.  Reported beside `value`, never as it."""
    import numpy as np

    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    g = HydroGrid(prm, ic, device=0)
    g.run(3, mode=nv.BX_STEP_GRAPH, endtime=1e300)
    g.sync()
    t0 = time.perf_counter()
    g.upload(ic)
    g.sync()
    t1 = time.perf_counter()
    g.run(steps, mode=nv.BX_STEP_GRAPH, endtime=1e300)
    g.sync()
    t2 = time.perf_counter()
    g.download()
    t3 = time.perf_counter()
    g.close()
    up_b = sum(8 * np.asarray(ic[k]).size for k in ("cons", "u", "p", "adindx") if ic.get(k) is not None)
    down_b = 8 * prm.ngrid * 9
    return {"workload": f"the headline box, {steps} graph-mode steps between an upload and a download",
            "upload_s": t1 - t0, "upload_GBps": up_b / (t1 - t0) / 1e9,
            "download_s": t3 - t2, "download_GBps": down_b / (t3 - t2) / 1e9,
            "steps_s": t2 - t1,
            "cell_updates_per_s_with_transfers": prm.interior_cells * steps / (t3 - t0),
            "note": "host arrays in the reference's layout (pageable memory); the library reorders on the device"}


def eos_cost(n: int, steps: int, reps: int):
    """The tabulated H2 EOS (H2STAT=0, the reference's default 10000 x 1000
    tables) on an n^3 molecular-cloud box: host table build, the state()
    kernel's HBM rate, and the exact-sequence step with tables against the
    same step with a fixed gamma."""
    import dataclasses

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = dataclasses.replace(box_params(n), H2STAT=0, gammafix=1.4, tk_bgrnd=10.0)
    ic = ics.h2_cloud_ic(prm, seed=0)
    t0 = time.perf_counter()
    g = HydroGrid(prm, ic, device=0)  # bx_eos_init builds the tables on the host and uploads them
    t_init = time.perf_counter() - t0
    for _ in range(3):
        g.state()
    g.sync()
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(reps):
            g.state()
        g.sync()
        best = min(best, (time.perf_counter() - t0) / reps)
    ncell = prm.ngrid
    interior = prm.interior_cells
    g.run(1, mode=nv.BX_STEP_EXACT)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, mode=nv.BX_STEP_EXACT)
    g.sync()
    t_tab = (time.perf_counter() - t0) / steps
    g.close()
    prm_fix = dataclasses.replace(prm, H2STAT=-1)
    g = HydroGrid(prm_fix, ic, device=0)
    g.run(1, mode=nv.BX_STEP_EXACT)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, mode=nv.BX_STEP_EXACT)
    g.sync()
    t_fix = (time.perf_counter() - t0) / steps
    g.close()
    bw = B_STATE_TAB * ncell / best
    return {"workload": f"H2 cloud box {n}^3 interior, H2STAT=0, tables 10000 T x 1000 rho (parameters.f90 "
                        f"sizes, tk_bgrnd 10 K), exact call sequence",
            "table_build_and_upload_s": t_init,
            "state_kernel": {"ms": best * 1e3, "cells": ncell, "alg_bytes_per_cell": B_STATE_TAB,
                             "achieved_GBps": bw / 1e9, "hbm_frac": bw / HBM_PEAK,
                             "note": "wall time of back-to-back bx_state calls (one kernel each) / calls"},
            "exact_step_ms_tabulated": t_tab * 1e3, "exact_step_ms_fixed_gamma": t_fix * 1e3,
            "tabulated_cell_updates_per_s": interior / t_tab,
            "tabulated_over_fixed_step_time": t_tab / t_fix}


def options_cost(n: int, steps: int):
    """Cost of flux_ang and of gravity from a given potential on an n^3
    rotating box: exact-sequence step times against the plain exact step,
    and the mass-moment reduction of external_grav in both orders."""
    import dataclasses

    import numpy as np

    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = box_params(n)
    ic = ics.rotating_box_ic(prm, seed=0, omega=0.002)
    x, y, z = ics.cell_centres(prm)
    phi = (-1.0 / np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2 + 100.0)).reshape(-1)

    def step_ms(p, grav=0, mode=nv.BX_STEP_EXACT, k=steps):
        g = HydroGrid(p, ic, device=0)
        if grav:
            g.set_gravity(grav, phi, object_mass=1.0, object_radius=20.0)
        g.run(1 if mode == nv.BX_STEP_EXACT else 4, mode=mode)
        g.sync()
        f0, _ = g.step_counts()
        t0 = time.perf_counter()
        g.run(k, mode=mode)
        g.sync()
        ms = (time.perf_counter() - t0) / k * 1e3
        if mode != nv.BX_STEP_EXACT:  # every timed step took the fused kernels
            f1, _ = g.step_counts()
            assert f1 - f0 == k and g.fallback_count() == 0, (f1 - f0, k)
            g.close()
            return ms, None
        moments = None
        if grav & nv.BX_GRAV_EXTERNAL:
            g.external_grav()
            t0 = time.perf_counter()
            for _ in range(3):
                g.external_grav()
            moments = (time.perf_counter() - t0) / 3 * 1e3
        g.close()
        return ms, moments

    base, _ = step_ms(prm)
    ang, _ = step_ms(dataclasses.replace(prm, fluxangmom=True))
    tree, tree_call = step_ms(prm, nv.BX_GRAV_POTENTIAL | nv.BX_GRAV_EXTERNAL)
    seq, seq_call = step_ms(prm, nv.BX_GRAV_POTENTIAL | nv.BX_GRAV_EXTERNAL | nv.BX_GRAV_COM_SEQUENTIAL)
    # a potential alone: the fused kernels (bit-identical to the exact sequence)
    pot, _ = step_ms(prm, nv.BX_GRAV_POTENTIAL)
    kf = 4 * steps
    pot_fused, _ = step_ms(prm, nv.BX_GRAV_POTENTIAL, nv.BX_STEP_FUSED, kf)
    pot_graph, _ = step_ms(prm, nv.BX_GRAV_POTENTIAL, nv.BX_STEP_GRAPH, kf)
    plain_graph, _ = step_ms(prm, 0, nv.BX_STEP_GRAPH, kf)
    return {"workload": f"rotating box {n}^3 interior FP64, exact call sequence, {steps} steps each",
            "exact_step_ms": base, "flux_ang_step_ms": ang, "flux_ang_over_plain": ang / base,
            "gravity_tree_step_ms": tree, "gravity_sequential_step_ms": seq,
            "external_grav_call_ms": {"tree": tree_call, "sequential": seq_call},
            "potential": {"exact_step_ms": pot, "fused_step_ms": pot_fused, "graph_step_ms": pot_graph,
                          "plain_graph_step_ms": plain_graph, "graph_over_plain_graph": pot_graph / plain_graph,
                          "exact_over_graph": pot / pot_graph,
                          "note": f"BX_GRAV_POTENTIAL alone, {kf} fused steps after 4 untimed; fused = one step "
                                  "per host round trip, graph = captured step pairs; plain = the same box "
                                  "without gravity"},
            "note": "gravity = calc_gforce + external_grav before source(tphase,1) each step; the sequential "
                    "mass moments (bit-identical to a one-thread reference) run on one wave"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--mode", choices=["fused", "exact"], default="fused")
    ap.add_argument("--tile", type=int, nargs=3, default=None, metavar=("YCHUNK", "ZCHUNK", "VARIANT"))
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the physical cores of the affinity mask")
    ap.add_argument("--mixed-size", type=int, default=384, help="box of the mixed-form line (config: 384)")
    ap.add_argument("--mixed-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched shock tubes")
    ap.add_argument("--no-boxes", action="store_true", help="skip the batched uneven boxes")
    ap.add_argument("--boxes-only", action="store_true", help="run only the batched uneven boxes")
    ap.add_argument("--graph-only", action="store_true", help="run only the graph-replay comparison")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint/forward cost")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay comparison")
    ap.add_argument("--boxes-walls", action="store_true",
                    help="with --boxes-only, also the same boxes with random walls (exact sequence only)")
    ap.add_argument("--no-boxes-seq", action="store_true",
                    help="batched boxes without the one-grid-at-a-time comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost")
    ap.add_argument("--no-eos", action="store_true", help="skip the tabulated-EOS line")
    ap.add_argument("--no-walls", action="store_true", help="skip the all-walls box line")
    ap.add_argument("--no-host", action="store_true", help="skip the line with host transfers")
    ap.add_argument("--no-options", action="store_true", help="skip the flux_ang / gravity line")
    ap.add_argument("--eos-only", action="store_true", help="run only the tabulated-EOS line")
    ap.add_argument("--options-only", action="store_true", help="run only the flux_ang / gravity line")
    ap.add_argument("--adjoint-size", type=int, default=128)
    ap.add_argument("--adjoint-large", type=int, default=256,
                    help="a second adjoint line at this size, where the segment stores cannot hold every "
                         "segment (0: none)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
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

    if args.eos_only:
        print(json.dumps({"eos": eos_cost(args.size, 5, 20)}))
        return 0
    if args.options_only:
        print(json.dumps({"options": options_cost(args.size, 3)}))
        return 0
    if args.adjoint_only:
        res = {"adjoint": adjoint_cost(args.adjoint_size, args.adjoint_steps, 20)}
        if args.adjoint_large:
            res["adjoint_256"] = adjoint_cost(args.adjoint_large, 200, 20)
        print(json.dumps(res))
        return 0
    if args.graph_only:
        print(json.dumps({"graph": graph_cost(600, 50, args.size)}))
        return 0
    if args.boxes_only:
        out = {"boxes": boxes_cost(512, 3, not args.no_boxes_seq)}
        if args.boxes_walls:
            out["boxes_walls"] = boxes_cost(512, 3, not args.no_boxes_seq, walls=True)
        print(json.dumps(out))
        return 0
    prm = box_params(args.size)
    log(f"generating box {args.size}^3 ...")
    ic = ics.box_ic(prm, seed=0)
    if args.mixed_study:
        print(json.dumps(mixed_study(prm, ic, args.steps)))
        return 0
    cells = prm.interior_cells
    if args.mode == "exact":
        g = HydroGrid(prm, ic, device=0)
        g.run(args.warmup, mode=nv.BX_STEP_EXACT)
        g.sync()
        t0 = time.perf_counter()
        g.run(args.steps, mode=nv.BX_STEP_EXACT)
        g.sync()
        el = time.perf_counter() - t0
        print(json.dumps({"metric": "exact-sequence step", "value": cells * args.steps / el,
                          "unit": "cell-updates/s", "ms_per_step": el * 1e3 / args.steps}))
        return 0

    log(f"FP64: warmup {args.warmup}, timing {args.steps} steps ...")
    value, ms_step, per, nprof, fb = timed_run(prm, ic, nv.BX_PREC_FP64, args.steps, args.warmup, args.tile)
    out = {
        "metric": "cell-updates/s @256^3 FP64 vs mixed (HBM roofline %); adjoint/forward cost",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D box {args.size}^3 interior FP64, outflow, gamma=5/3, CFL 0.5, "
                               f"seeded smoothed Gaussian rho/p, rms Mach 0.5; full main-loop step",
                   "grid": [prm.nx, prm.ny, prm.nz], "mode": "fused, replayed from hipGraphs (BX_STEP_GRAPH)",
                   "fallback_steps": fb},
        "step_hbm_roofline_frac": value * B_STEP_FP64 / HBM_PEAK,
    }
    variant = args.tile[2] if args.tile else 11
    kname = ("bx::k_stage<2, double, %d> (corrector stage: x/y/z faces, update, next Courant rate)"
             % (8 if variant == 11 else 4)) if variant >= 10 else \
        "bx::k_xz<2, double, ...> (corrector stage, x/z faces + update)"
    k2 = per[3] * 1e-3
    kernel_alg = B_XZ2_FP64 * cells  # U*, U^n read, U^{n+1} written: 5 FP64 fields each
    traffic = None
    tfile = os.path.join("profiles", f"traffic_fused_step_box{args.size}.json")
    if os.path.exists(os.path.join(ROOT, tfile)):
        with open(os.path.join(ROOT, tfile)) as f:
            tj = json.load(f)
        traffic = {"step_bytes": tj.get("step_hbm_bytes"), "kernel_bytes": tj.get("corrector_hbm_bytes"),
                   "source": f"{tfile}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes of this workload, "
                             f"scaled by a plain copy's counters (measured there, not in this run)"}
    step_achieved = value * B_STEP_FP64
    out["roofline"] = {"scope": "the main-loop step (both stage kernels, the ghost fills, the dt kernel), "
                                "replayed from hipGraphs; 200 algorithmic bytes per cell-update",
                       "bound": "hbm", "achieved": step_achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": step_achieved / HBM_PEAK,
                       "traffic": traffic["step_bytes"] if traffic else None,
                       "traffic_source": traffic["source"] if traffic else None,
                       "alg_bytes_per_step": B_STEP_FP64 * cells,
                       "kernel": {"name": kname, "achieved": kernel_alg / k2 / 1e9, "unit": "GB/s",
                                  "frac": kernel_alg / k2 / HBM_PEAK, "alg_bytes_per_launch": kernel_alg,
                                  "avg_launch_ms": per[3], "traffic": traffic["kernel_bytes"] if traffic else None,
                                  "timing": "HIP events on the library's stream around each launch",
                                  "launches_timed": nprof},
                       "stage_kernel_avg_ms": {"stage1": per[0] + per[1], "stage2": per[2] + per[3]}}
    if not args.no_walls:
        log("all-walls box ...")
        out["walls"] = walls_cost(args.size, args.steps, args.warmup, ms_step)
    if not args.no_host:
        log("with host transfers ...")
        out["with_host_transfers"] = host_cost(prm, ic, args.steps)
    if not args.no_mixed:
        msz = args.mixed_size
        log(f"mixed form at {msz}^3: {args.mixed_steps} steps each precision ...")
        mprm = box_params(msz)
        mic = ic if msz == args.size else ics.box_ic(mprm, seed=0)
        out["mixed"] = mixed_study(mprm, mic, args.mixed_steps)
        del mic
    if not args.no_graph:
        log("graph replays ...")
        out["graph"] = graph_cost(600, args.steps, args.size)
    if not args.no_batch:
        log("batched shock tubes ...")
        out["batched"] = batched(512, 3)
    if not args.no_boxes:
        log("batched uneven boxes ...")
        out["boxes"] = boxes_cost(512, 3, not args.no_boxes_seq)
    if not args.no_eos:
        log("tabulated EOS ...")
        out["eos"] = eos_cost(args.size, 5, 20)
    if not args.no_options:
        log("flux_ang and gravity ...")
        out["options"] = options_cost(args.size, 3)
    if not args.no_adjoint:
        log(f"adjoint {args.adjoint_size}^3 x {args.adjoint_steps} ...")
        out["adjoint"] = adjoint_cost(args.adjoint_size, args.adjoint_steps, 20)
        if args.adjoint_large:
            log(f"adjoint {args.adjoint_large}^3 x 200 ...")
            out["adjoint_256"] = adjoint_cost(args.adjoint_large, 200, 20)
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_steps} steps, {args.cpu_threads} threads) ...")
        out["cpu_baseline"] = cpu_baseline(prm, ic, args.cpu_steps, args.cpu_threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
