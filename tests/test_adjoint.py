"""Adjoint of the FP64 step (bx_adjoint_*, boxzyhydro_amd.adjoint) against
central finite differences of the GPU forward with the same frozen dt
sequence: per-cell gradient entries and directional derivatives (the
forward's tangent), plus determinism and the autograd wiring."""
import numpy as np
import pytest

from boxzyhydro_amd import ics
from boxzyhydro_amd.params import box_params

EPS = 1e-6


def _setup(n=16, nsteps=8, every=3, seed=0):
    import torch

    from boxzyhydro_amd.adjoint import AdjointStepper

    prm = box_params(n)
    ic = ics.box_ic(prm, seed=seed)
    tic = ics.box_ic(prm, seed=seed + 1)
    dev = torch.device("cuda", 0)
    st = AdjointStepper(prm, ic["u"], nsteps, every)
    rho0 = torch.as_tensor(ic["cons"][:, 0].copy(), device=dev)
    p0 = torch.as_tensor(ic["p"].copy(), device=dev)
    # target: the second seed evolved forward with its own Courant dt
    tst = AdjointStepper(prm, tic["u"], nsteps, 0)
    tgt = tst.forward(torch.as_tensor(tic["cons"][:, 0].copy(), device=dev),
                      torch.as_tensor(tic["p"].copy(), device=dev), checkpoint=False)
    interior = torch.as_tensor(ics.boundary_flags(prm).reshape(-1) == 0, device=dev)
    st.forward(rho0, p0)  # fixes the dt sequence
    return prm, st, rho0, p0, tgt, interior


def _final(st, rho, p):
    return st.forward(rho, p, dts=st.dts, checkpoint=False)


def _dloss(a, b, tgt, interior):
    """L(a) - L(b) for L = |rho_T - tgt|^2/2 over interior cells, term by term
    ((a-b)(a+b-2 tgt)/2: the difference of the final densities is formed
    before any cancellation)."""
    return float((0.5 * (a - b) * (a + b - 2 * tgt) * interior).sum())


def _central(st, rho0, p0, d_rho, d_p, eps, tgt, interior):
    a = _final(st, rho0 + eps * d_rho, p0 + eps * d_p)
    b = _final(st, rho0 - eps * d_rho, p0 - eps * d_p)
    base = _final(st, rho0, p0)
    return (_dloss(a, b, tgt, interior) / (2 * eps), _dloss(a, base, tgt, interior) / eps,
            _dloss(base, b, tgt, interior) / eps)


def _grad(st, rho0, p0, tgt, interior):
    from boxzyhydro_amd.adjoint import hydro_function

    f = hydro_function()
    rho = rho0.clone().requires_grad_(True)
    p = p0.clone().requires_grad_(True)
    rT = f.apply(rho, p, st)
    d = (rT - tgt) * interior
    L = 0.5 * (d * d).sum()
    L.backward()
    return rho.grad.detach(), p.grad.detach(), float(L.detach())


@pytest.mark.gpu
def test_adjoint_gradient_matches_central_differences():
    """64 seeded interior gradient entries (32 cells, rho and p) of a 16^3 box
    over 8 steps with checkpoints every 3 (a partial last segment), against
    central differences with eps = 1e-6: |fd - adjoint| <= 1e-6 max|g|.

    The step is piecewise smooth (limiter argument, HLL speed bounds), so a
    +-eps stencil can straddle a branch switch; such an entry is accepted only
    when the two one-sided differences disagree (a kink inside the stencil)
    and the central difference agrees at a smaller eps -- at most 4 of 64."""
    import torch

    prm, st, rho0, p0, tgt, interior = _setup()
    g_rho, g_p, L = _grad(st, rho0, p0, tgt, interior)
    rng = np.random.Generator(np.random.PCG64(7))
    cells = rng.choice(np.flatnonzero(interior.cpu().numpy()), size=32, replace=False)
    scale = max(float(g_rho.abs().max()), float(g_p.abs().max()))
    tol = 1e-6 * scale
    zero = torch.zeros_like(rho0)
    kinks, worst = [], 0.0
    for c in cells:
        for which, g in (("rho", g_rho), ("p", g_p)):
            e = torch.zeros_like(rho0)
            e[c] = 1.0
            d_rho, d_p = (e, zero) if which == "rho" else (zero, e)
            ad = float(g[c])
            fd, fwd, bwd = _central(st, rho0, p0, d_rho, d_p, EPS, tgt, interior)
            err = abs(fd - ad)
            print(f"cell {c} {which}: adjoint {ad:+.12e} fd {fd:+.12e} err/max|g| {err / scale:.2e}", flush=True)
            if err <= tol:
                worst = max(worst, err / scale)
                continue
            # a branch switch inside the stencil: the one-sided differences disagree
            assert abs(fwd - bwd) > 10 * tol, (c, which, fd, fwd, bwd, ad)
            refined = min(abs(_central(st, rho0, p0, d_rho, d_p, h, tgt, interior)[0] - ad) for h in (1e-7, 1e-8))
            print(f"   kink inside +-eps (one-sided {fwd:+.6e} / {bwd:+.6e}); refined err/max|g| "
                  f"{refined / scale:.2e}", flush=True)
            assert refined <= tol, (c, which)
            kinks.append((int(c), which))
    print(f"worst smooth entry err/max|g| {worst:.2e}; entries with a kink inside +-1e-6: {kinks}", flush=True)
    assert len(kinks) <= 4


@pytest.mark.gpu
def test_adjoint_matches_forward_tangent():
    """<grad, v> against the forward's directional derivative along random
    directions v in (rho, p), taken by central differences as eps shrinks
    (a random direction crosses branch switches for larger eps)."""
    import torch

    prm, st, rho0, p0, tgt, interior = _setup(nsteps=10, every=4, seed=3)
    g_rho, g_p, _ = _grad(st, rho0, p0, tgt, interior)
    gen = torch.Generator(device="cpu").manual_seed(11)
    for k in range(3):
        vr = torch.randn(rho0.shape, generator=gen, dtype=torch.float64).to(rho0.device) * interior
        vp = torch.randn(p0.shape, generator=gen, dtype=torch.float64).to(p0.device) * interior
        ad = float((g_rho * vr).sum() + (g_p * vp).sum())
        errs = []
        for h in (1e-7, 1e-8, 1e-9):
            fd = _central(st, rho0, p0, vr, vp, h, tgt, interior)[0]
            errs.append(abs(ad - fd) / abs(fd))
            print(f"direction {k} eps {h:.0e}: adjoint {ad:+.12e} tangent {fd:+.12e} rel {errs[-1]:.2e}", flush=True)
        assert min(errs) <= 1e-6


@pytest.mark.gpu
def test_adjoint_deterministic_and_checkpoint_independent():
    """The gradient does not depend on the checkpoint spacing and repeats bit
    for bit."""
    import torch

    prm, st, rho0, p0, tgt, interior = _setup(nsteps=7, every=7)
    g1 = _grad(st, rho0, p0, tgt, interior)
    g2 = _grad(st, rho0, p0, tgt, interior)
    assert torch.equal(g1[0], g2[0]) and torch.equal(g1[1], g2[1])
    st.every = 2
    g3 = _grad(st, rho0, p0, tgt, interior)
    assert torch.equal(g1[0], g3[0]) and torch.equal(g1[1], g3[1])
    # ghost cells carry no gradient (their values are overwritten by the fill)
    assert float(g1[0][~interior].abs().max()) == 0.0


@pytest.mark.gpu
def test_adjoint_courant_forward_matches_step():
    """The adjoint's forward with its own Courant dts (formed on the device,
    read back once) is bx_step's fused step from the same regular state: the
    same dt sequence and final state, bit for bit."""
    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.hydro import HydroGrid

    prm, st, rho0, p0, tgt, interior = _setup(nsteps=9, every=4)
    d_adj = np.asarray(st.dts).copy()
    final = st.state().cpu().numpy().T.copy()  # (n, 5)
    # the same regular state in a context of its own: ghosts filled, u = m/rho
    cons = st.cons_of(rho0, p0).cpu().numpy().T.copy()
    g = HydroGrid(prm, {"cons": cons, "u": cons[:, 1:4] / cons[:, 0:1], "p": p0.cpu().numpy().copy(),
                        "adindx": np.full(prm.ngrid, prm.gammafix)})
    g.set_ghost_cells()
    c = g.download()["cons"]
    g.upload({"cons": c, "u": c[:, 1:4] / c[:, 0:1], "p": p0.cpu().numpy().copy(),
              "adindx": np.full(prm.ngrid, prm.gammafix)})
    d_step = g.run(9, endtime=1e300, mode=nv.BX_STEP_FUSED)
    assert np.array_equal(d_adj.view(np.uint64), np.asarray(d_step).view(np.uint64))
    assert np.array_equal(final.view(np.uint64), g.download()["cons"].view(np.uint64))
    g.close()


@pytest.mark.gpu
def test_adjoint_kept_and_recomputed_segments_agree():
    """The gradient is the same bits whether the forward keeps every segment's
    states (no recompute), keeps two, or keeps one (each other segment
    recomputed from its checkpoint, the store reused), and over two backward
    sweeps of one forward."""
    import torch

    prm, st, rho0, p0, tgt, interior = _setup(nsteps=11, every=3)  # 4 segments, a partial last one
    seg_bytes = prm.ngrid * 8 * (5 * (3 + 1) + 6 * 3)
    st.store(-1)
    g_all = _grad(st, rho0, p0, tgt, interior)
    r0 = st.store(-1)
    assert r0 == 0  # every segment kept: nothing recomputed
    for nbytes, want in ((2 * seg_bytes, 2), (0, 3)):
        st.store(nbytes)
        g = _grad(st, rho0, p0, tgt, interior)
        r1 = st.store(nbytes)
        assert r1 - r0 == want, (nbytes, r1 - r0)
        r0 = r1
        assert torch.equal(g[0], g_all[0]) and torch.equal(g[1], g_all[1]), nbytes
    # a second sweep of the same forward (one store: every segment recomputed)
    lam = (st.forward(rho0, p0, dts=st.dts) - tgt) * interior
    a = st.backward(lam)
    b = st.backward(lam)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert st.store(0) - r0 == 3 + 4


@pytest.mark.gpu
def test_adjoint_full_size_matches_central_differences():
    """The adjoint workload at its full size: 128^3 box, 200 steps, checkpoints
    every 20, loss 0.5|rho_T - rho_target|^2 with the target from a second
    seed.  64 seeded interior cells, rho and p each, against central
    differences with eps = 1e-6: |fd - adjoint| <= 1e-6 max|g|.  Over 200
    steps a +-eps stencil can straddle a branch switch of the limiter or the
    HLL speed bounds somewhere in the cell's domain of dependence: such an
    entry passes only when its one-sided differences disagree (a kink inside
    the stencil) and a smaller eps agrees -- at most 4 of 128."""
    import torch

    prm, st, rho0, p0, tgt, interior = _setup(n=128, nsteps=200, every=20)
    g_rho, g_p, _ = _grad(st, rho0, p0, tgt, interior)
    rng = np.random.Generator(np.random.PCG64(11))
    cells = rng.choice(np.flatnonzero(interior.cpu().numpy()), size=64, replace=False)
    scale = max(float(g_rho.abs().max()), float(g_p.abs().max()))
    tol = 1e-6 * scale
    zero = torch.zeros_like(rho0)
    kinks, errs, rels = [], [], []
    print(f"128^3 x 200 steps, max|g| {scale:.4e}", flush=True)
    for c in cells:
        for which, g in (("rho", g_rho), ("p", g_p)):
            e = torch.zeros_like(rho0)
            e[c] = 1.0
            d_rho, d_p = (e, zero) if which == "rho" else (zero, e)
            ad = float(g[c])
            fd, fwd, bwd = _central(st, rho0, p0, d_rho, d_p, EPS, tgt, interior)
            err = abs(fd - ad)
            if err <= tol:
                errs.append(err / scale)
                rels.append((err / max(abs(ad), 1e-300), abs(ad) / scale))
                continue
            print(f"cell {c} {which}: adjoint {ad:+.12e} fd {fd:+.12e} err/max|g| {err / scale:.2e} "
                  f"one-sided gap/max|g| {abs(fwd - bwd) / scale:.2e}", flush=True)
            assert abs(fwd - bwd) > tol, (c, which, fd, fwd, bwd, ad)
            refined = min(abs(_central(st, rho0, p0, d_rho, d_p, h, tgt, interior)[0] - ad) for h in (1e-7, 1e-8))
            print(f"   refined err/max|g| {refined / scale:.2e}", flush=True)
            assert refined <= tol, (c, which)
            kinks.append((int(c), which))
    print(f"smooth entries: {len(errs)}, worst err/max|g| {max(errs):.2e}, median {np.median(errs):.2e}; "
          f"kinks {kinks}", flush=True)
    # per entry, relative to the entry itself: the FD error is O(eps^2) in the
    # third derivative plus O(ulp |L| / eps) of rounding, so entries far below
    # max|g| carry a larger relative error; reported by size class
    r = np.array([x for x, _ in rels])
    mag = np.array([m for _, m in rels])
    print(f"relative |fd-g|/|g| over {len(r)} smooth entries: median {np.median(r):.2e}, "
          f"90% {np.quantile(r, 0.9):.2e}, max {r.max():.2e}; above 1e-6: {int((r > 1e-6).sum())}, "
          f"above 1e-4: {int((r > 1e-4).sum())}", flush=True)
    for lo, hi in ((1e-1, 2.0), (1e-2, 1e-1), (1e-3, 1e-2), (0.0, 1e-3)):
        sel = (mag >= lo) & (mag < hi)
        if sel.any():
            print(f"   |g|/max|g| in [{lo:g}, {hi:g}): {int(sel.sum())} entries, max relative "
                  f"{r[sel].max():.2e}, above 1e-6: {int((r[sel] > 1e-6).sum())}", flush=True)
    assert len(kinks) <= 4
    # entries within a decade of the largest are checked relative to themselves
    big = mag >= 0.1
    assert (r[big] <= 1e-6).all(), r[big].max()


@pytest.mark.gpu
def test_adjoint_refused_forward_leaves_nothing_to_sweep():
    """The backward needs a completed checkpointed forward: before any
    forward, after one that keeps no checkpoints, and after one refused for a
    state that is not regular (a density below 1e-4), it returns
    BX_ERR_STATE instead of sweeping an earlier run."""
    import torch

    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.adjoint import AdjointStepper

    prm = box_params(12)
    ic = ics.box_ic(prm, seed=4)
    dev = torch.device("cuda", 0)
    rho0 = torch.as_tensor(ic["cons"][:, 0].copy(), device=dev)
    p0 = torch.as_tensor(ic["p"].copy(), device=dev)
    g = torch.ones_like(rho0)

    def refused(call, code):
        with pytest.raises(nv.BoxzyError) as e:
            call()
        assert f"({code})" in str(e.value)

    st = AdjointStepper(prm, ic["u"], 4, 2)
    refused(lambda: st.backward(g), nv.BX_ERR_STATE)
    st.forward(rho0, p0)
    g_ok = st.backward(g)
    st.forward(rho0, p0, checkpoint=False)
    refused(lambda: st.backward(g), nv.BX_ERR_STATE)
    st.forward(rho0, p0)
    assert torch.equal(st.backward(g)[0], g_ok[0])
    low = rho0.clone()
    interior = np.flatnonzero(ics.boundary_flags(prm).reshape(-1) == 0)
    low[int(interior[len(interior) // 2])] = 5e-5
    refused(lambda: st.forward(low, p0), nv.BX_ERR_UNSUPPORTED)
    refused(lambda: st.backward(g), nv.BX_ERR_STATE)


@pytest.mark.gpu
def test_adjoint_two_rollouts_backward_of_the_first():
    """Two applications on one stepper, then the backward of the first
    output: the context holds the second forward's checkpoints, so the first
    node runs its own forward again (from its saved inputs and dt sequence)
    and returns the same bits as a single application."""
    import torch

    from boxzyhydro_amd.adjoint import hydro_function

    prm, st, rho0, p0, tgt, interior = _setup(n=16, nsteps=8, every=3, seed=2)
    f = hydro_function()

    def loss_grad(apply_second):
        rho = rho0.clone().requires_grad_(True)
        p = p0.clone().requires_grad_(True)
        r1 = f.apply(rho, p, st)
        if apply_second:
            r2 = f.apply(rho0 * 1.01, p0 * 0.99, st)  # another rollout, not part of the loss
            assert not torch.equal(r1, r2)
        loss = (0.5 * (r1 - tgt) ** 2 * interior).sum()
        loss.backward()
        return rho.grad.clone(), p.grad.clone()

    g_single = loss_grad(False)
    g_two = loss_grad(True)
    assert torch.equal(g_single[0], g_two[0]) and torch.equal(g_single[1], g_two[1])


@pytest.mark.gpu
def test_adjoint_refusal_mid_run_restores_the_start():
    """A forward refused after its steps ran (a diverging flow whose density
    falls below 1e-4 within the run) puts the starting state back (its
    ghosts filled by the forward's first call) and marks u, p and the Courant
    rate stale: a download matches that state, and fused steps from it match
    a fresh context started from the same state."""
    import ctypes

    import torch

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
    g = HydroGrid(prm, ic)
    rc = nv.lib().bx_adjoint_forward(g._h, 40, 10, None, None)
    assert rc == nv.BX_ERR_UNSUPPORTED and b"rho<1e-4" in nv.lib().bx_last_error()
    want = HydroGrid(prm, ic)
    want.set_ghost_cells()  # the forward's own first call
    w = want.download()["cons"]
    out = g.download()
    assert np.array_equal(out["cons"].view(np.uint64), w.view(np.uint64))
    with np.errstate(divide="ignore", invalid="ignore"):
        assert np.array_equal(out["u"].view(np.uint64), (out["cons"][:, 1:4] / out["cons"][:, :1]).view(np.uint64))
    s0 = torch.empty(5 * g.ngrid, dtype=torch.float64, device="cuda")
    nv.check(nv.lib().bx_cons_device(g._h, ctypes.c_void_p(s0.data_ptr()), 0), "bx_cons_device")
    fresh = HydroGrid(prm, None)
    nv.check(nv.lib().bx_cons_device(fresh._h, ctypes.c_void_p(s0.data_ptr()), 1), "bx_cons_device")
    d1 = g.run(20, endtime=1e30, mode=nv.BX_STEP_FUSED)
    d2 = fresh.run(20, endtime=1e30, mode=nv.BX_STEP_FUSED)
    assert np.array_equal(d1.view(np.uint64), d2.view(np.uint64))
    o1, o2 = g.download(), fresh.download()
    for k in ("cons", "u", "p"):
        assert np.array_equal(o1[k].view(np.uint64), o2[k].view(np.uint64)), k
    for h in (g, want, fresh):
        h.close()


@pytest.mark.gpu
def test_adjoint_refuses_corner_ghost_boxes():
    """A 3-D box walled in y and z has two-level corner ghosts, whose fix-up
    after the corrector is not differentiated: the forward is refused."""
    import dataclasses

    import torch

    from boxzyhydro_amd import native as nv
    from boxzyhydro_amd.adjoint import AdjointStepper

    prm = dataclasses.replace(box_params(12), no_outflow_yl=True, no_outflow_zl=True)
    ic = ics.box_ic(prm, seed=4)
    dev = torch.device("cuda", 0)
    st = AdjointStepper(prm, ic["u"], 4, 2)
    with pytest.raises(nv.BoxzyError, match="corner ghosts"):
        st.forward(torch.as_tensor(ic["cons"][:, 0].copy(), device=dev), torch.as_tensor(ic["p"].copy(), device=dev))
