"""bx_step(BX_STEP_GRAPH): runs of fused steps replayed from a captured hipGraph
with dt formed on the device.  The same bits as BX_STEP_FUSED and as the
reference:

* every golden case (the reference's own runs, tests/golden/golden.json): the
  dt sequence and the final cons/u/p.  sod_t02 ends on its endtime (the last
  steps are taken one at a time);
* a diverging flow whose density falls below 1e-4 inside a replayed run: the
  run is undone and retaken step by step, with the same bits as the fused
  mode;
* a 64^3 box in FP64 and in the mixed form against BX_STEP_FUSED, over runs
  that start on either buffer of the step pair;
* host-only: the chunk bound keeps every replayed step below endtime.
"""
import numpy as np
import pytest

from cases import case_inputs, sha

CASES = ["sod_step1", "box12_outflow", "box12_walls", "box24_outflow", "stress8", "flatz_walls",
         "riemann4_0", "riemann4_1", "riemann4_2", "riemann4_3", "sod_t02", "readme_sod_y"]
# long outflow runs: every ghost donor is an interior cell, so the fused path
# (and with it the replays) applies; grids with walls chain ghost donors and
# take the exact sequence in every mode
REPLAYED = ["riemann4_0", "riemann4_1", "riemann4_2", "riemann4_3", "sod_t02"]


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_graph_steps_match_reference(golden, name):
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = golden[name]
    prm, ic = case_inputs(rec)
    print(f"case {name} graph", flush=True)
    g = HydroGrid(prm, ic)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_GRAPH)
    assert list(dts) == rec["dt"]
    out = g.download()
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    print(f"  replayed steps {g.graph_step_count()}, fallbacks {g.fallback_count()}", flush=True)
    if name in REPLAYED:
        assert g.graph_step_count() > 0
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_graph_equals_fused_box64(precision):
    """70 steps in calls of 3, 40 and 27 steps: the first call leaves the
    state on the other buffer of the pair (an odd step count), so both
    captured pairs are used."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = box_params(64)
    ic = ics.box_ic(prm, seed=3)
    prec = nv.BX_PREC_FP64 if precision == "fp64" else nv.BX_PREC_MIXED
    res = {}
    for mode in (nv.BX_STEP_FUSED, nv.BX_STEP_GRAPH):
        print(f"box 64^3 {precision} mode {mode}", flush=True)
        g = HydroGrid(prm, ic, precision=prec)
        d = np.concatenate([g.run(n, endtime=1e30, mode=mode) for n in (3, 40, 27)])
        res[mode] = (d, g.download(), g.time, g.dtold, g.step_count, g.graph_step_count())
        g.close()
    (df, of, tf, dof, sf, _), (dg, og, tg, dog, sg, ng) = res[nv.BX_STEP_FUSED], res[nv.BX_STEP_GRAPH]
    assert _bits_equal(df, dg)
    assert (tf, dof, sf) == (tg, dog, sg)
    for k in ("cons", "u", "p"):
        assert _bits_equal(of[k], og[k]), k
    assert ng >= 60


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_graph_run_undone_on_low_density(precision):
    """A diverging flow empties the middle of an outflow box: the density
    there first drops below 1e-4 (the low-density branch of state.f90:52-56,
    which the fused kernels leave to the exact sequence) after step 26.  Steps
    2-19 of the first call are replayed; the second call's run (steps 21-50)
    meets the branch, is undone and retaken one step at a time.  Both modes
    give the same bits and the same exact-sequence fallbacks."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=44, ny=12, nz=12, dx=1.0, dy=1.0, dz=1.0, gammafix=1.4, cfl=0.5, endtime=1e30,
                      no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False, no_outflow_yr=False,
                      no_outflow_zl=False, no_outflow_zr=False)
    x, _, _ = ics.cell_centres(prm)
    shp = prm.shape
    rho0 = 0.006
    vx = np.broadcast_to(np.sign(x)[None, None, :] * 1.0, shp)
    ic = ics._assemble(prm, np.full(shp, rho0), vx, np.zeros(shp), np.zeros(shp), np.full(shp, 0.01 * rho0))
    prec = nv.BX_PREC_FP64 if precision == "fp64" else nv.BX_PREC_MIXED
    res = {}
    for mode in (nv.BX_STEP_FUSED, nv.BX_STEP_GRAPH):
        g = HydroGrid(prm, ic, precision=prec)
        d = np.concatenate([g.run(20, mode=mode), g.run(30, mode=mode)])
        res[mode] = (d, g.download(), g.fallback_count(), g.graph_step_count(), g.time)
        print(f"mode {mode}: fallbacks {res[mode][2]}, replayed {res[mode][3]}", flush=True)
        g.close()
    (df, of, ff, _, tf), (dg, og, fg, ng, tg) = res[nv.BX_STEP_FUSED], res[nv.BX_STEP_GRAPH]
    assert len(df) == 50 and _bits_equal(df, dg) and tf == tg
    for k in ("cons", "u", "p"):
        assert _bits_equal(of[k], og[k]), k
    assert ff == fg and ff > 0
    assert ng == 18


@pytest.mark.gpu
def test_gpu_graph_after_tune_and_upload():
    """Captured pairs are dropped by bx_tune (new tile shapes) and recaptured;
    an upload keeps the buffers, so the pairs stay valid.  The same calls in
    fused mode give the same bits."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    prm = box_params(40)
    ic1, ic2 = ics.box_ic(prm, seed=11), ics.box_ic(prm, seed=12)
    res = {}
    for mode in (nv.BX_STEP_FUSED, nv.BX_STEP_GRAPH):
        g = HydroGrid(prm, ic1)
        d = [g.run(15, endtime=1e30, mode=mode)]
        g.tune(8, 8, 0)
        d.append(g.run(15, endtime=1e30, mode=mode))
        g.upload(ic2)
        g.time, g.dtold, g.step_count = 0.0, 0.0, 0
        d.append(g.run(15, endtime=1e30, mode=mode))
        res[mode] = (np.concatenate(d), g.download()["cons"], g.graph_step_count())
        g.close()
    (df, cf, _), (dg, cg, ng) = res[nv.BX_STEP_FUSED], res[nv.BX_STEP_GRAPH]
    assert _bits_equal(df, dg) and _bits_equal(cf, cg)
    assert ng >= 36


def test_chunk_bound_keeps_steps_below_endtime():
    """The host bound behind the replayed runs (mirrored here): with each dt
    at most 1.1x the last, K steps from (time, dtold) stay below endtime."""
    rng = np.random.Generator(np.random.PCG64(0))
    for _ in range(2000):
        time = float(rng.uniform(0, 1))
        dtold = float(10 ** rng.uniform(-6, -1))
        endtime = time + float(10 ** rng.uniform(-6, 1))
        K, t, d = 0, time, dtold
        while K < 32:
            d = d * 1.1 * (1 + 1e-9)
            t = (t + d) * (1 + 1e-9)
            if not t < endtime:
                break
            K += 1
        K &= ~1
        # worst case: every dt the largest the cap allows, rounded the C way
        t, d = time, dtold
        for _ in range(K):
            d = d * 1.1
            assert d + t <= endtime  # no clip (main.f90:276)
            t = t + d
            assert t < endtime or K == 0
