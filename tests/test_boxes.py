"""Batched grids (bx_boxes_*): many independent grids of uneven size in one
persistent launch, each bit-identical to the reference run of that grid.

* The reference's own results (tests/golden/golden.json, made with
  oracle/_ref): every golden case -- 3-D boxes with and without walls, the
  branch-forcing stress box, a flat-z box with walls, the Sod tubes and the
  seeded Riemann problems -- in ONE batch; each case's dt sequence and final
  cons/u/p are bit-identical.
* Uneven random grids (3-D and flat-z, random walls, spacings, gamma, CFL
  number) against the one-kernel-per-subroutine path (HydroGrid, exact mode),
  bit for bit, for every work split, with and without the captured graph,
  and on a replay.
* Argument checks (host only).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from cases import case_inputs, sha, stress_ic

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def _final_time(rec):
    t = rec["time"]
    return float(t[-1] if isinstance(t, list) else t)


def _golden_batch():
    """Every golden case, run to its recorded final time: the last step then
    meets dt + time == endtime exactly (no clip), so each case takes its
    recorded number of steps."""
    names = sorted(GOLDEN)
    grids = [case_inputs(GOLDEN[n]) for n in names]
    return names, grids, [_final_time(GOLDEN[n]) for n in names]


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, False])
def test_gpu_golden_cases_in_one_batch_match_reference(graph):
    from boxzyhydro_amd.boxes import GridBatch

    names, grids, ends = _golden_batch()
    print(f"boxes: {len(names)} golden cases, graph={graph}", flush=True)
    b = GridBatch(grids, ends)
    most = max(GOLDEN[n]["steps"] for n in names)
    b.log_dt(most)
    b.run(maxsteps=most + 1, graph=graph)
    b.sync()
    res = b.result()
    dts = b.dts()
    for i, name in enumerate(names):
        rec = GOLDEN[name]
        r = res[i]
        assert r["status"] == 0, name
        assert r["steps"] == rec["steps"], name
        assert [float(x) for x in dts[i]] == rec["dt"], name
        for k in ("cons", "u", "p"):
            assert sha(r[k]) == rec["final_sha256"][k], (name, k)
    b.close()


def _random_grids(seed, count):
    from boxzyhydro_amd import ics
    from boxzyhydro_amd.params import HydroParams

    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        flat = rng.random() < 0.35
        nx, ny = (int(v) for v in rng.integers(5, 19, 2))
        nz = int(rng.integers(1, 6)) if flat else int(rng.integers(6, 15))
        walls = [bool(w) for w in rng.integers(0, 2, 6)]
        prm = HydroParams(nx=nx, ny=ny, nz=nz, dx=float(rng.choice([0.25, 0.1, 0.3])),
                          dy=float(rng.choice([0.25, 0.2, 0.125])), dz=float(rng.choice([0.25, 0.15])),
                          no_outflow_xl=walls[0], no_outflow_xr=walls[1], no_outflow_yl=walls[2],
                          no_outflow_yr=walls[3], no_outflow_zl=walls[4], no_outflow_zr=walls[5],
                          gammafix=float(rng.choice([1.4, 5.0 / 3.0, 1.3])), cfl=float(rng.uniform(0.2, 0.8)),
                          endtime=1e30)
        if i % 5 == 4 and min(nx, ny) >= 8 and nz >= 8:
            ic = stress_ic(prm, seed=1000 + i)  # low density, vlimit clamp and energy floors
        else:
            ic = ics.box_ic(prm, seed=2000 + i, sigma_cells=float(rng.uniform(1.0, 3.0)),
                            mach_rms=float(rng.uniform(0.2, 1.2)))
        out.append((prm, ic))
    return out


def _exact_runs(grids, steps):
    """Each grid through the one-kernel-per-subroutine path: (dts, final time, arrays)."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    ref = []
    for i, (prm, ic) in enumerate(grids):
        g = HydroGrid(prm, ic, device=0)
        d = g.run(int(steps[i]), mode=nv.BX_STEP_EXACT)
        out = g.download()
        ref.append((list(d), g.time, {k: out[k] for k in ("cons", "u", "p")}))
        g.close()
    return ref


@pytest.mark.gpu
def test_gpu_uneven_random_grids_match_exact_path():
    from boxzyhydro_amd.boxes import GridBatch

    grids = _random_grids(seed=7, count=30)
    steps = np.random.Generator(np.random.PCG64(8)).integers(1, 12, len(grids))
    print(f"boxes: {len(grids)} random grids, {sum(p.ngrid for p, _ in grids)} cells", flush=True)
    ref = _exact_runs(grids, steps)
    b = GridBatch(grids, [t for _, t, _ in ref])
    b.log_dt(int(steps.max()))
    for split in ((16, 1), (8, 1), (4, 1), (4, 2)):
        for graph in (True, False):
            print(f"boxes: split {split} graph {graph}", flush=True)
            b.tune(*split)
            for _ in range(2):  # a replay starts again from the set-ups
                b.run(maxsteps=int(steps.max()) + 1, graph=graph)
                b.sync()
                res, dts = b.result(), b.dts()
                for i, (d, t, arrs) in enumerate(ref):
                    assert res[i]["steps"] == steps[i], (i, split)
                    assert res[i]["time"] == t, (i, split)
                    assert [float(x) for x in dts[i]] == d, (i, split)
                    for k, v in arrs.items():
                        assert np.array_equal(res[i][k].view(np.uint64), v.view(np.uint64)), (i, split, k)
    b.close()


@pytest.mark.gpu
def test_gpu_endtime_clip_and_queue_order():
    """Grids of one set-up with different endtimes: the last step of each is
    clipped to its endtime (main.f90:276-278) as in the exact path, and the
    queue puts the longest run first."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.boxes import GridBatch
    from boxzyhydro_amd.hydro import HydroGrid

    prm, ic = _random_grids(seed=11, count=1)[0]
    ends = [0.05, 0.2, 0.1]
    b = GridBatch([(prm, ic)] * 3, ends)
    assert list(b.order()) == [1, 2, 0]
    b.run(maxsteps=1000)
    b.sync()
    res = b.result()
    for i, e in enumerate(ends):
        g = HydroGrid(prm, ic, device=0)
        g.run(1000, endtime=e, mode=nv.BX_STEP_EXACT)
        out = g.download()
        assert res[i]["time"] == g.time == e
        assert res[i]["steps"] == g.step_count
        assert np.array_equal(res[i]["cons"].view(np.uint64), out["cons"].view(np.uint64))
        g.close()
    b.close()


@pytest.mark.gpu
def test_gpu_dt_log_on_off_between_graph_runs():
    """bx_boxes_log_dt between replays of the captured launch: the log is
    turned off (dts refused) and on again with another length, the launch is
    recaptured each time, and every run returns the same bits."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.boxes import GridBatch

    grids = _random_grids(seed=31, count=3)
    b = GridBatch(grids)
    states, logs = [], []
    for maxlog in (6, 0, 3):
        b.log_dt(maxlog)
        b.run(maxsteps=6, graph=True)
        b.sync()
        res = b.result()
        assert all(r["steps"] == 6 for r in res)
        states.append([r["cons"].copy() for r in res])
        if maxlog:
            dts = b.dts()
            assert [len(d) for d in dts] == [maxlog] * 3
            logs.append(dts)
        else:
            with pytest.raises(nv.BoxzyError):
                b.dts()
    for run in states[1:]:
        for x, y in zip(states[0], run):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    for x, y in zip(logs[0], logs[1]):
        assert np.array_equal(x[:3].view(np.uint64), y.view(np.uint64))
    b.close()


@pytest.mark.gpu
def test_gpu_many_small_grids_stop_at_maxsteps():
    """600 tiny grids (more than the persistent launch's workgroups, so the
    queue hands out work many times over) stopped by maxsteps before their
    endtime: every grid takes exactly that many steps, and a sample equals the
    exact path's run of as many steps."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.boxes import GridBatch
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import HydroParams

    rng = np.random.Generator(np.random.PCG64(21))
    grids = []
    for i in range(600):
        nx, ny = (int(v) for v in rng.integers(5, 9, 2))
        nz = int(rng.integers(1, 7))
        prm = HydroParams(nx=nx, ny=ny, nz=nz, dx=0.25, dy=0.25, dz=0.25, gammafix=1.4, cfl=0.5, endtime=1e30)
        grids.append((prm, ics.box_ic(prm, seed=5000 + i, sigma_cells=1.0)))
    b = GridBatch(grids)
    b.run(maxsteps=4)
    b.sync()
    res = b.result()
    assert all(r["steps"] == 4 and r["status"] == 0 for r in res)
    for i in range(0, 600, 60):
        prm, ic = grids[i]
        g = HydroGrid(prm, ic, device=0)
        g.run(4, mode=nv.BX_STEP_EXACT)
        out = g.download()
        assert res[i]["time"] == g.time, i
        for k in ("cons", "u", "p"):
            assert np.array_equal(res[i][k].view(np.uint64), out[k].view(np.uint64)), (i, k)
        g.close()
    b.close()


def test_create_rejects_unsupported_grids():
    """Checked on the host before any device call."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import make_params
    from boxzyhydro_amd.params import HydroParams

    def create(prms):
        arr = (nv.BxParams * len(prms))(*[make_params(p) for p in prms])
        n = sum(p.ngrid for p in prms)
        cons, u, p, ends = np.ones((n, 5)), np.zeros((n, 3)), np.ones(n), np.ones(len(prms))
        h = ctypes.c_void_p()
        return nv.lib().bx_boxes_create(len(prms), arr, nv.dptr(ends), nv.dptr(cons), nv.dptr(u), nv.dptr(p), 0,
                                        ctypes.byref(h))

    ok = HydroParams(nx=8, ny=8, nz=8)
    for bad in (HydroParams(nx=4, ny=8, nz=8), HydroParams(nx=8, ny=8, nz=8, fluxangmom=True),
                HydroParams(nx=8, ny=8, nz=8, dx=0.0), HydroParams(nx=8, ny=8, nz=8, gammafix=1.0)):
        rc = create([ok, bad])
        assert rc == nv.BX_ERR_ARG
        assert "grid 1" in nv.lib().bx_last_error().decode()
    arr = (nv.BxParams * 1)(make_params(ok))
    h = ctypes.c_void_p()
    assert nv.lib().bx_boxes_create(0, arr, None, None, None, None, 0, ctypes.byref(h)) == nv.BX_ERR_ARG


def test_batch_rejects_nonzero_starttime():
    from boxzyhydro_amd.boxes import GridBatch
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=8, ny=8, nz=8, starttime=1.0)
    with pytest.raises(ValueError):
        GridBatch([(prm, {"cons": np.ones((prm.ngrid, 5)), "u": np.zeros((prm.ngrid, 3)), "p": np.ones(prm.ngrid)})])


def test_batch_checks_each_setup_and_the_endtimes():
    """Sizes are checked per grid (two mismatched set-ups whose total is right
    are refused) and there must be one endtime per grid; both before any call
    into the library."""
    from boxzyhydro_amd.boxes import GridBatch
    from boxzyhydro_amd.params import HydroParams

    a, b = HydroParams(nx=8, ny=8, nz=8), HydroParams(nx=9, ny=8, nz=8)

    def setup(prm, extra=0):
        n = prm.ngrid + extra
        return {"cons": np.ones((n, 5)), "u": np.zeros((n, 3)), "p": np.ones(n)}

    with pytest.raises(ValueError, match="grid 0"):
        GridBatch([(a, setup(a, -8)), (b, setup(b, 8))])
    with pytest.raises(ValueError, match="endtimes"):
        GridBatch([(a, setup(a)), (b, setup(b))], endtimes=[0.1])
