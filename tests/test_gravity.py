"""Gravity from a given potential: calc_gforce (utils.f90:549-571), the
external point mass (external_grav, gravity.f90:1477-1549) and the source
term with per-cell forces (source.f90:148-182), against the reference's own
code (golden vectors in tests/golden/grav.json, made by
tests/golden/make_grav_golden.py with oracle/_ref) and closed forms."""
import hashlib
import json
import os

import numpy as np
import pytest

from cases import Driver, HipAdapter, case_inputs, sha

HERE = os.path.dirname(os.path.abspath(__file__))
GRAV = json.load(open(os.path.join(HERE, "golden", "grav.json")))


def _flags(rec, tree=False):
    from boxzyhydro_amd import native as nv

    f = nv.BX_GRAV_POTENTIAL | (nv.BX_GRAV_EXTERNAL if rec["grav"] == 2 else 0)
    return f | (0 if tree else nv.BX_GRAV_COM_SEQUENTIAL)


def _grid(rec, tree=False):
    from boxzyhydro_amd.hydro import HydroGrid

    prm, ic = case_inputs(rec)
    phi = ic.pop("phi")
    g = HydroGrid(prm, ic, device=0)
    g.set_gravity(_flags(rec, tree), phi)  # with the calls of main.f90:146-154
    return prm, g


@pytest.mark.parametrize("name", sorted(GRAV))
def test_generator_reproduces_golden_inputs(name):
    rec = GRAV[name]
    _, ic = case_inputs(rec)
    assert hashlib.sha256(np.ascontiguousarray(ic["phi"], "<f8").tobytes()).hexdigest() == rec["phi_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "fused", "graph"])
@pytest.mark.parametrize("name", sorted(GRAV))
def test_gpu_steps_match_reference(name, mode):
    """Sequential mass moments: bit-identical to the reference on one thread.
    With a potential alone the fused and graph modes take the fused kernels
    after the first step (from the set-up); with the point mass, whose force
    needs the corrector's mass moments, they run the exact sequence."""
    from boxzyhydro_amd import native as nv

    rec = GRAV[name]
    print(f"gravity case {name} {mode}", flush=True)
    _, g = _grid(rec)
    m = {"exact": nv.BX_STEP_EXACT, "fused": nv.BX_STEP_FUSED, "graph": nv.BX_STEP_GRAPH}[mode]
    dts = g.run(rec["steps"], mode=m)
    assert [float(x) for x in dts] == rec["dt"]
    out = g.download()
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    fused, exact = g.step_counts()
    print(f"  fused steps {fused}, exact steps {exact}, fallbacks {g.fallback_count()}", flush=True)
    if mode == "exact" or rec["grav"] == 2:
        assert fused == 0
    else:
        assert exact == 1 and g.fallback_count() == 0 and fused == rec["steps"] - 1
        if mode == "graph" and rec["steps"] >= 5:
            assert g.graph_step_count() > 0
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "graph"])
def test_gpu_fused_potential_change_mid_run(mode):
    """A new phi between calls: the next step's source(0) still takes the
    force of the old phi and its calc_gforce that of the new one, so that step
    runs the exact sequence; the fused kernels resume after it.  Bit-identical
    to the exact mode making the same calls."""
    from boxzyhydro_amd import native as nv

    rec = GRAV["grav_pot_xwalls"]
    prm, ic = case_inputs(rec)
    phi2 = 0.7 * ic["phi"] + 0.05 * np.cos(np.arange(ic["phi"].size) * 0.01)
    outs = {}
    for m in ("exact", mode):
        print(f"gravity phi change {m}", flush=True)
        _, g = _grid(rec)
        step = nv.BX_STEP_EXACT if m == "exact" else (nv.BX_STEP_FUSED if m == "fused" else nv.BX_STEP_GRAPH)
        d1 = g.run(4, mode=step)
        nv.check(nv.lib().bx_set_phi(g._h, nv.dptr(np.ascontiguousarray(phi2))), "bx_set_phi")
        d2 = g.run(6, mode=step)
        out = g.download()
        outs[m] = (list(d1) + list(d2), {k: sha(out[k]) for k in ("cons", "u", "p")})
        if m != "exact":
            fused, exact = g.step_counts()
            assert exact == 2 and fused == 8 and g.fallback_count() == 0, (fused, exact)
        g.close()
    assert outs["exact"] == outs[mode]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(n for n in GRAV if "snaps" in GRAV[n]))
def test_gpu_subroutines_match_reference_snapshots(name):
    rec = GRAV[name]
    print(f"gravity snapshots {name}", flush=True)
    prm, g = _grid(rec)
    ad = HipAdapter(g)
    got = {}

    def snap(tag, dt):
        got[tag] = (dt, {k: sha(v) for k, v in ad.arrays().items()})

    dts, _ = Driver(ad).run(rec["steps"], prm.endtime, snap=snap)
    for s in rec["snaps"]:
        if s["tag"] == 0:
            continue
        dt, digests = got[s["tag"]]
        assert dt == s["dt"], f"snapshot {s['tag']}: dt"
        for k, v in s["sha256"].items():
            assert digests[k] == v, f"snapshot {s['tag']}: {k}"
    assert list(dts) == rec["dt"]


@pytest.mark.gpu
def test_gpu_tree_moments_close_to_sequential():
    """Without BX_GRAV_COM_SEQUENTIAL the mass moments are summed in a fixed
    tree order instead of cell order: the moments agree to rounding, and so do
    the runs."""
    from boxzyhydro_amd import native as nv

    rec = GRAV["grav_ext_disk"]
    _, gs = _grid(rec)
    _, gt = _grid(rec, tree=True)
    ms, mt = gs.external_grav(), gt.external_grav()
    np.testing.assert_allclose(mt, ms, rtol=1e-13, atol=1e-15)
    ds = gs.run(rec["steps"], mode=nv.BX_STEP_EXACT)
    dtr = gt.run(rec["steps"], mode=nv.BX_STEP_EXACT)
    np.testing.assert_allclose(dtr, ds, rtol=1e-11)
    np.testing.assert_allclose(gt.download()["cons"], gs.download()["cons"], rtol=1e-10, atol=1e-13)


@pytest.mark.gpu
def test_gpu_calc_gforce_linear_potential_is_exact():
    """phi = a x + b y + c z on dyadic spacings: the centred differences are
    exact, so gforce = -(a, b, c) on interior cells and stays 0 on ghosts
    (init_grid.f90:299)."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=12, ny=10, nz=9, dx=0.5, dy=0.25, dz=0.5, endtime=1e30)
    x, y, z = ics.cell_centres(prm)
    a, b, c = 3.0, -2.0, 0.75
    phi = (a * x[None, None, :] + b * y[None, :, None] + c * z[:, None, None]).reshape(-1)
    g = HydroGrid(prm, ics.box_ic(prm, seed=0), device=0)
    g.set_gravity(nv.BX_GRAV_POTENTIAL, phi)
    gf = g.download_gforce()
    inner = ics.boundary_flags(prm).reshape(-1) == 0
    assert np.all(gf[inner] == [-a, -b, -c])
    assert np.all(gf[~inner] == 0.0)


@pytest.mark.gpu
def test_gpu_external_point_mass_far_field():
    """Beyond twice the softening radius the spline is 1 and its derivative 0
    (utils.f90:486-510): the added force is the point mass's m/r^2 towards
    (xobj, yobj) = -(moments)/m."""
    from boxzyhydro_amd import ics
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=24, ny=20, nz=1, dx=0.1, dy=0.1, dz=0.1, endtime=1e30)
    g = HydroGrid(prm, ics.box_ic(prm, seed=5), device=0)
    mass, radius = 2.5, 0.3
    g.set_gravity(nv.BX_GRAV_EXTERNAL, None, mass, radius)  # phi = 0: no calc_gforce term
    com = g.external_grav()  # second call: the force doubles
    gf = g.download_gforce()
    x, y, _ = ics.cell_centres(prm)
    X = np.broadcast_to(x[None, :], (prm.ny, prm.nx)).reshape(-1) + com[0] / mass
    Y = np.broadcast_to(y[:, None], (prm.ny, prm.nx)).reshape(-1) + com[1] / mass
    r = np.sqrt(X * X + Y * Y)
    far = r > 2.0 * radius
    np.testing.assert_allclose(gf[far, 0], 2 * (-mass * X[far] / r[far] ** 3), rtol=1e-13)
    np.testing.assert_allclose(gf[far, 1], 2 * (-mass * Y[far] / r[far] ** 3), rtol=1e-13)
    assert np.all(gf[:, 2] == 0.0)


@pytest.mark.gpu
def test_gpu_tuned_variant_with_gravity():
    """Only the tiled 64 x 8 stage kernel has the gravity form: configuring a
    potential returns an earlier tuned variant to the default split, bx_tune
    then refuses the others, and the run stays bit-identical to the
    reference."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    rec = GRAV["grav_pot_xwalls"]
    prm, ic = case_inputs(rec)
    phi = ic.pop("phi")
    print("gravity with a tuned variant", flush=True)
    g = HydroGrid(prm, ic, device=0)
    g.tune(8, 16, 1)  # a two-kernel variant
    g.set_gravity(nv.BX_GRAV_POTENTIAL, phi)
    with pytest.raises(nv.BoxzyError, match="variant 11"):
        g.tune(8, 16, 4)
    g.tune(1, 8, 11)
    dts = g.run(rec["steps"], mode=nv.BX_STEP_FUSED)
    assert [float(x) for x in dts] == rec["dt"]
    out = g.download()
    for k in ("cons", "u", "p"):
        assert sha(out[k]) == rec["final_sha256"][k], k
    assert g.step_counts() == (rec["steps"] - 1, 1) and g.fallback_count() == 0
    g.close()
