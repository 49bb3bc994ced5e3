"""Batched shock tubes (bx_batch_*): the row form against the reference's
golden runs and the oracle, bit for bit.

CPU tests pin the premise of the row form on the oracle (a tube's full grid
stays row-uniform through every branch of the step) and the host-side checks;
GPU tests run the persistent launch, plain and as a replayed hipGraph."""
import numpy as np
import pytest

from boxzyhydro_amd import ics
from cases import case_inputs, oracle_tube, sha, stress_tubes

GOLDEN_TUBES = ["riemann4_0", "riemann4_1", "riemann4_2", "riemann4_3"]


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def _same(a, b):
    return np.array_equal(_bits(a), _bits(b))


# ----------------------------------------------------------------- CPU side
def test_tube_rows_match_full_setups():
    """riemann_tubes builds exactly the rows of riemann_batch's 3-D set-ups."""
    from boxzyhydro_amd.batch import grid_of_row, tube_row_of

    full = ics.riemann_batch(8, seed=3, lmin=64, lmax=512)
    rows = ics.riemann_tubes(8, seed=3, lmin=64, lmax=512)
    for (p1, ic), (p2, row) in zip(full, rows):
        assert p1 == p2
        got = tube_row_of(p1, ic)
        for k in ("cons", "u", "p"):
            assert _same(got[k], row[k]), k
        back = grid_of_row(p2, row)
        for k in ("cons", "u", "p", "adindx"):
            assert _same(back[k], ic[k]), k


def test_tube_row_of_rejects_non_tubes():
    from boxzyhydro_amd.batch import tube_row_of

    prm, ic = ics.riemann_batch(1, seed=0, lmin=64, lmax=64)[0]
    bad = {k: v.copy() for k, v in ic.items()}
    bad["cons"][prm.nx * 3 + 10, 0] *= 1.5  # one cell of row iy=4
    with pytest.raises(ValueError):
        tube_row_of(prm, bad)


@pytest.mark.parametrize("i", range(6))
def test_row_form_premise_on_oracle(i):
    """The oracle's full-grid run of each stress tube is row-uniform after
    every step taken (asserted inside oracle_tube), through the low-density
    EOS branch, the vlimit clamp, non-cubic cells and a power-of-two dx."""
    prm, row = stress_tubes()[i]
    dts, times, fin, grid = oracle_tube(prm, row)
    if i == 0:
        assert (grid["cons"][:, 0] < 1e-4).any()  # the low-density branch is met
    if i == 1:
        assert np.abs(grid["u"][:, 0]).max() == pytest.approx(prm.vlimit)  # the clamp acts
    if i == 5:
        assert len(dts) == 0


def test_batch_check_rejects_non_tubes():
    """bx_batch_check (host only) names the tube that breaks the row form."""
    import dataclasses

    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.batch import check_tubes

    tubes = stress_tubes()[:3]
    check_tubes(tubes)
    prm, row = tubes[1]
    cases = [
        (dataclasses.replace(prm, nz=6), row, "flat z"),
        (dataclasses.replace(prm, no_outflow_xl=True), row, "reflecting"),
        (prm, {**row, "cons": np.where(np.arange(prm.nx)[:, None] == 7, 5e-5, row["cons"])}, "density"),
        (prm, {**row, "u": np.where(np.arange(prm.nx)[:, None] == 1, 0.25, row["u"])}, "velocity"),
    ]
    for p2, r2, what in cases:
        with pytest.raises(nv.BoxzyError, match="tube 1"):
            check_tubes([tubes[0], (p2, r2), tubes[2]])


# ----------------------------------------------------------------- GPU side
def _batch(tubes, **kw):
    from boxzyhydro_amd.batch import TubeBatch

    return TubeBatch(tubes, **kw)


def test_batch_checks_rows_and_endtimes():
    """Each tube's arrays must be one x-row of its grid, and there must be one
    endtime per tube: refused on the host before any call into the library."""
    from boxzyhydro_amd.batch import TubeBatch, check_tubes

    tubes = ics.riemann_tubes(2, seed=3, lmin=64, lmax=96)
    (pa, ra), (pb, rb) = tubes
    short = {k: v[:-1] for k, v in ra.items()}
    with pytest.raises(ValueError, match="tube 0"):
        check_tubes([(pa, short), (pb, rb)])
    with pytest.raises(ValueError, match="tube 0"):
        TubeBatch([(pa, short), (pb, rb)])
    with pytest.raises(ValueError, match="endtimes"):
        TubeBatch(tubes, endtimes=[0.1, 0.1, 0.1])


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_reference_goldens(golden, graph):
    """The four golden tubes (reference Fortran runs of the full 3-D grids):
    dt sequences bit-exact, final cons/u/p of the expanded rows hash-equal."""
    from boxzyhydro_amd.batch import expand_row, tube_row_of

    tubes, recs = [], []
    for name in GOLDEN_TUBES:
        rec = golden[name]
        prm, ic = case_inputs(rec)
        tubes.append((prm, tube_row_of(prm, ic)))
        recs.append(rec)
    print(f"golden tubes graph={graph}", flush=True)
    b = _batch(tubes)
    b.log_dt(max(r["steps"] for r in recs))
    b.run(graph=graph)
    b.sync()
    res = b.result()
    dts = b.dts()
    for (prm, _), rec, r, d in zip(tubes, recs, res, dts):
        assert r["steps"] == rec["steps"]
        assert list(d) == rec["dt"]
        assert r["time"] == rec["time"][-1]
        full = expand_row(prm, r)
        for k in ("cons", "u", "p"):
            assert sha(full[k]) == rec["final_sha256"][k], k


@pytest.mark.gpu
def test_batch_matches_oracle_stress_tubes():
    """Low-density branch, clamp, gamma=5/3 with non-cubic cells, dx=2^-6, the
    shortest tube and a tube with no step: rows, dt, time bitwise vs the oracle."""
    from boxzyhydro_amd.batch import STATUS_CLAMP, STATUS_LOWRHO

    tubes = stress_tubes()
    print("stress tubes", [p.nx for p, _ in tubes], flush=True)
    b = _batch(tubes)
    b.log_dt(4096)
    b.run(graph=False)
    b.sync()
    res = b.result()
    dts = b.dts()
    for i, ((prm, row), r, d) in enumerate(zip(tubes, res, dts)):
        odts, otimes, fin, _ = oracle_tube(prm, row)
        assert r["steps"] == len(odts), i
        assert _same(d, odts), i
        if len(odts):
            assert r["time"] == otimes[-1]
        for k in ("cons", "u", "p"):
            assert _same(r[k], fin[k]), (i, k)
    assert res[0]["status"] & STATUS_LOWRHO
    assert res[1]["status"] & STATUS_CLAMP
    assert res[3]["status"] == 0


@pytest.mark.gpu
def test_batch_random_tubes_match_oracle():
    """Eight tubes of the 512-problem generator (seed 0, the short ones, full
    endtime) against the oracle, while running inside the whole 512 batch."""
    allt = ics.riemann_tubes(512, seed=0)
    pick = [i for i, (p, _) in enumerate(allt) if p.nx <= 200][:8]
    print("random tubes", pick, flush=True)
    b = _batch(allt)
    b.run(graph=True)
    ms = b.sync()
    res = b.result()
    print(f"  512 tubes in {ms:.1f} ms", flush=True)
    assert all(r["status"] & 4 == 0 for r in res)
    assert all(r["time"] >= 0.1 for r in res)
    for i in pick:
        prm, row = allt[i]
        odts, otimes, fin, _ = oracle_tube(prm, row, threads=8)
        assert res[i]["steps"] == len(odts)
        assert res[i]["time"] == otimes[-1]
        for k in ("cons", "u", "p"):
            assert _same(res[i][k], fin[k]), (i, k)


@pytest.mark.gpu
def test_batch_graph_replays_and_tiles_agree():
    """Replays of the captured graph, with other device work between them, and
    every work split of the launch give the same rows as a plain launch."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid
    from boxzyhydro_amd.params import box_params

    tubes = ics.riemann_tubes(48, seed=5, lmin=64, lmax=1024)
    b = _batch(tubes)
    b.run(graph=False)
    b.sync()
    ref = b.result()
    other = HydroGrid(box_params(16), ics.box_ic(box_params(16), seed=1))
    for rep in range(3):
        print(f"replay {rep}", flush=True)
        b.run(graph=True)
        b.sync()
        got = b.result()
        for r0, r1 in zip(ref, got):
            assert r0["steps"] == r1["steps"]
            for k in ("cons", "u", "p"):
                assert _same(r0[k], r1[k])
        other.run(2, mode=nv.BX_STEP_FUSED)  # unrelated work between replays
        other.sync()
    for waves, bpc in ((4, 1), (8, 2), (16, 1), (4, 4)):
        print(f"tile {waves} {bpc}", flush=True)
        b.tune(waves, bpc)
        b.run(graph=True)
        b.sync()
        got = b.result()
        for r0, r1 in zip(ref, got):
            for k in ("cons", "u", "p"):
                assert _same(r0[k], r1[k])


@pytest.mark.gpu
def test_batch_result_independent_of_batch_and_order():
    """A tube's result does not depend on which other tubes share the launch."""
    allt = ics.riemann_tubes(64, seed=9, lmin=64, lmax=2048)
    b = _batch(allt)
    b.run()
    b.sync()
    full = b.result()
    sub = [allt[i] for i in (63, 7, 30)]
    b2 = _batch(sub)
    b2.run(graph=False)
    b2.sync()
    for j, i in enumerate((63, 7, 30)):
        r = b2.result()[j]
        for k in ("cons", "u", "p"):
            assert _same(r[k], full[i][k])


def _long_tubes():
    import json
    import os

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tubes_long.json")) as f:
        return json.load(f)


def test_long_tube_goldens_match_generator():
    """The three long-tube goldens (tests/golden/make_tube_golden.py: the
    tubes that set the 512-tube batch's run time, reference runs of their full
    grids) were made from this generator's set-ups."""
    from boxzyhydro_amd.batch import grid_of_row
    from cases import ic_digest

    recs = _long_tubes()
    allt = ics.riemann_tubes(512, seed=0)
    assert sorted(r["generator"]["index"] for r in recs.values()) == [26, 156, 241]
    for rec in recs.values():
        prm, row = allt[rec["generator"]["index"]]
        assert ic_digest(grid_of_row(prm, row)) == rec["ic_sha256"]
        assert len(rec["dt_sampled"]) == (rec["steps"] + rec["dt_every"] - 1) // rec["dt_every"]


@pytest.mark.gpu
@pytest.mark.parametrize("gang", [(0, 0), (4, 32)])
def test_batch_long_tubes_match_reference_goldens(gang):
    """Inside the whole 512-tube batch, the three tubes that set its run time
    (up to 4,049 cells and 16,891 steps) against the reference's runs of their
    full grids: step count, every dt (SHA-256 of the sequence, and every 64th
    value), final time, and the final cons/u/p of the expanded rows.  With
    gangs, the 32 costliest tubes (these three among them) each run over 4
    workgroups."""
    import hashlib

    from boxzyhydro_amd.batch import expand_row

    recs = _long_tubes()
    allt = ics.riemann_tubes(512, seed=0)
    b = _batch(allt)
    b.log_dt(max(r["steps"] for r in recs.values()))
    print(f"512 tubes, long-tube goldens, gangs {gang}", flush=True)
    if gang[0]:
        b.run(graph=True)
        b.sync()  # measures every tube's steps, which pick the gang tubes
        b.gang(*gang)
    b.run(graph=True)
    b.sync()
    assert b.gang_fallbacks() == 0
    res = b.result()
    dts = b.dts()
    for name, rec in recs.items():
        i = rec["generator"]["index"]
        prm, _ = allt[i]
        r, d = res[i], np.asarray(dts[i], np.float64)
        print(f"  {name}: {r['steps']} steps", flush=True)
        assert r["steps"] == rec["steps"], name
        assert list(d[::rec["dt_every"]]) == rec["dt_sampled"], name
        assert hashlib.sha256(d.tobytes()).hexdigest() == rec["dt_sha256"], name
        assert r["time"] == rec["time_final"], name
        full = expand_row(prm, r)
        for k in ("cons", "u", "p"):
            assert sha(full[k]) == rec["final_sha256"][k], (name, k)


@pytest.mark.gpu
@pytest.mark.parametrize("gang", [(2, 6), (3, 4), (4, 8)])
def test_batch_gangs_match_one_workgroup_per_tube(gang):
    """Gangs (each of the costliest tubes over several workgroups, the cells at
    the shared edges exchanged through device-coherent accesses) give the bits
    of one workgroup per tube: final rows, time, steps, status and every dt."""
    tubes = ics.riemann_tubes(24, seed=5)
    b = _batch(tubes)
    b.log_dt(4000)
    b.run(graph=False)
    b.sync()
    ref, ref_dt = b.result(), b.dts()
    b.gang(*gang)
    print(f"24 tubes, gangs {gang}", flush=True)
    for graph in (False, True):
        b.run(graph=graph)
        b.sync()
        assert b.gang_fallbacks() == 0
        got, got_dt = b.result(), b.dts()
        for i, (x, y) in enumerate(zip(got, ref)):
            for k in ("cons", "u", "p"):
                assert _same(x[k], y[k]), (gang, graph, i, k)
            assert (x["time"], x["steps"], x["status"]) == (y["time"], y["steps"], y["status"]), (gang, i)
            assert _same(got_dt[i], ref_dt[i]), (gang, i)


@pytest.mark.gpu
def test_batch_gang_limits():
    """members x gangs must fit the compute units; members < 2 turns gangs off."""
    from boxzyhydro_amd import native as nv

    b = _batch(ics.riemann_tubes(8, seed=1))
    with pytest.raises(nv.BoxzyError):
        b.gang(64, 8)   # 512 workgroups: above the compute units
    with pytest.raises(nv.BoxzyError):
        b.gang(2, 9)    # more gangs than tubes
    b.gang(1, 4)        # off
    b.run(graph=False)
    b.sync()
    assert b.gang_fallbacks() == 0
