"""Differentiating the hydro step: the hand-written HIP adjoint of the FP64
step (bx_adjoint_* of include/boxzy.h) behind a ``torch.autograd.Function``.

The differentiated map takes the initial density and pressure (velocity held
fixed, energy rebuilt as in the set-up, E = p/(gamma-1) + rho |u|^2/2) through
``nsteps`` steps of main.f90:211-555 with a frozen dt sequence to the final
state.  The forward keeps a checkpoint every ``every`` steps and, for as many
of the last segments as its store budget holds (``AdjointStepper.store``;
default half the free device memory), the per-step states of the segment.
The backward sweeps the segments in reverse: kept segments as they are, the
others recomputed from their checkpoints first.  Both give the same bits.

torch is used for the tensors and the autograd plumbing only: the steps and
their adjoint run in libboxzy_hip.so on the context's own stream.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import native as nv
from .hydro import HydroGrid
from .params import HydroParams


def _torch():
    import torch

    return torch


class AdjointStepper:
    """A context that runs checkpointed forwards and their adjoints.

    ``u`` (ngrid, 3) is the fixed velocity field of the set-up.  Tensors are
    (ngrid,) float64 on the context's device, in the reference's cell order.
    """

    def __init__(self, prm: HydroParams, u: np.ndarray, nsteps: int, every: int = 20, device: int = 0):
        torch = _torch()
        self.prm = prm
        self.nsteps = nsteps
        self.every = every
        self.device = torch.device("cuda", device)
        self.grid = HydroGrid(prm, None, device=device)
        n = prm.ngrid
        self.n = n
        self.u = torch.as_tensor(np.ascontiguousarray(u, np.float64).T.copy(), device=self.device)  # (3, n)
        self.uu = self.u[0] * self.u[0] + self.u[1] * self.u[1] + self.u[2] * self.u[2]
        self.gm1 = prm.gammafix - 1.0
        self.dts: Optional[np.ndarray] = None
        # checkpointed forwards run so far; the context holds the checkpoints
        # of the last one only
        self.generation = 0

    # --------------------------------------------------------------- helpers
    def cons_of(self, rho, p):
        """(5, n) conserved planes of the set-up convention (ics._assemble)."""
        torch = _torch()
        c = torch.empty((5, self.n), dtype=torch.float64, device=self.device)
        c[0] = rho
        c[1] = rho * self.u[0]
        c[2] = rho * self.u[1]
        c[3] = rho * self.u[2]
        c[4] = p / self.gm1 + rho * 0.5 * self.uu
        return c

    def _set(self, planes) -> None:
        _torch().cuda.synchronize(self.device)
        nv.check(nv.lib().bx_cons_device(self.grid._h, ctypes.c_void_p(planes.data_ptr()), 1), "bx_cons_device")

    def state(self):
        """(5, n) conserved planes of the context's current state."""
        torch = _torch()
        out = torch.empty((5, self.n), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize(self.device)
        nv.check(nv.lib().bx_cons_device(self.grid._h, ctypes.c_void_p(out.data_ptr()), 0), "bx_cons_device")
        return out

    # ------------------------------------------------------------- forward
    def forward(self, rho, p, dts: Optional[np.ndarray] = None, checkpoint: bool = True):
        """Final density after nsteps steps from (rho, p).  With ``dts`` the
        steps use that frozen sequence; without, the Courant dt of each step,
        which is then kept as ``self.dts`` for later frozen runs."""
        self._set(self.cons_of(rho, p).contiguous())
        out = np.zeros(self.nsteps)
        dptr = nv.dptr(np.ascontiguousarray(dts, np.float64)) if dts is not None else nv.dptr(None)
        if checkpoint:
            self.generation += 1
        nv.check(nv.lib().bx_adjoint_forward(self.grid._h, self.nsteps, self.every if checkpoint else 0, dptr,
                                             nv.dptr(out)), "bx_adjoint_forward")
        if dts is None:
            self.dts = out
        return self.state()[0].clone()

    # ------------------------------------------------------------- backward
    def backward(self, grad_rho_T):
        """Adjoints of the initial (rho, p) for the adjoint ``grad_rho_T`` of
        the final density (of the last checkpointed forward)."""
        torch = _torch()
        lamT = torch.zeros((5, self.n), dtype=torch.float64, device=self.device)
        lamT[0] = grad_rho_T
        lam0 = torch.empty_like(lamT)
        torch.cuda.synchronize(self.device)
        nv.check(nv.lib().bx_adjoint_backward(self.grid._h, ctypes.c_void_p(lamT.data_ptr()),
                                              ctypes.c_void_p(lam0.data_ptr())), "bx_adjoint_backward")
        # E = p/(gamma-1) + rho |u|^2/2, m = rho u
        g_rho = lam0[0] + lam0[1] * self.u[0] + lam0[2] * self.u[1] + lam0[3] * self.u[2] + lam0[4] * 0.5 * self.uu
        g_p = lam0[4] / self.gm1
        return g_rho, g_p

    def tune(self, y_chunk: int, z_chunk: int) -> None:
        nv.check(nv.lib().bx_adjoint_tune(self.grid._h, y_chunk, z_chunk), "bx_adjoint_tune")

    def store(self, nbytes: int = -1) -> int:
        """Device memory the next checkpointed forward may spend on keeping
        whole segments (-1: half the free memory; 0: one segment, the others
        recomputed by the backward).  Returns the number of segments backward
        sweeps have recomputed so far."""
        r = ctypes.c_int64()
        nv.check(nv.lib().bx_adjoint_store(self.grid._h, nbytes, ctypes.byref(r)), "bx_adjoint_store")
        return r.value


def hydro_function():
    """The ``torch.autograd.Function`` (built lazily so that importing this
    module needs no torch)."""
    torch = _torch()

    class HydroSteps(torch.autograd.Function):
        """rho_T = steps(rho_0, p_0) on an :class:`AdjointStepper` with its
        frozen dt sequence (``stepper.dts``, or the Courant sequence of this
        forward when none is set)."""

        @staticmethod
        def forward(ctx, rho, p, stepper: AdjointStepper):
            ctx.stepper = stepper
            out = stepper.forward(rho.detach(), p.detach(), dts=stepper.dts, checkpoint=True)
            # the context keeps the checkpoints of its last forward only: a
            # later forward on the same stepper (another rollout in the same
            # graph) replaces them, and this node's backward then runs its
            # own forward again from the saved inputs
            ctx.generation = stepper.generation
            ctx.dts = np.array(stepper.dts, copy=True)
            ctx.save_for_backward(rho.detach(), p.detach())
            return out

        @staticmethod
        def backward(ctx, grad):
            st = ctx.stepper
            if st.generation != ctx.generation:
                rho, p = ctx.saved_tensors
                st.forward(rho, p, dts=ctx.dts, checkpoint=True)
            g_rho, g_p = st.backward(grad.contiguous())
            return g_rho, g_p, None

    return HydroSteps
