"""Differentiable level-set advection on top of the HIP kernels.

* rk3_step(blk, phi_n, ucn, dt): one ad_RKTVD3 step as a torch.autograd
  Function whose backward is the hand-written adjoint kernel set
  (lsm_rk3_step_adjoint): gradients with respect to phi^n and to U/J.
* advect(blk, phi0, ucn, dt, nsteps, checkpoint_every=k): nsteps steps with
  checkpointing: the forward keeps phi every k steps; the backward recomputes
  each segment from its checkpoint and runs the adjoint steps in reverse, so
  memory holds nsteps/k + k fields instead of nsteps.
* contravariant(blk, u, v, w): U/J = (1/aj) grad(xi_d) . u (rhs_contra_j.F90:
  28-48) in torch operations, so gradients reach the Cartesian velocity (and
  whatever parametrises it).

The adjoint differentiates the WENO3 nonlinear weights; the upwind choice, the
forced boundary stencils and the BigPhi switch of bcond_lsm are taken as in
the forward (a.e. derivatives of the piecewise-smooth step).
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from . import api


def contravariant(blk: "api.Block", u, v, w):
    """U_d / J at every node (the operation order of rhs_contra_j.F90:32-46)."""
    tmp = 1.0 / blk.aj
    U1 = tmp * ((blk.csi[0] * u + blk.csi[1] * v) + blk.csi[2] * w)
    U2 = tmp * ((blk.eta[0] * u + blk.eta[1] * v) + blk.eta[2] * w)
    U3 = tmp * ((blk.zet[0] * u + blk.zet[1] * v) + blk.zet[2] * w)
    return U1, U2, U3


class _RK3Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi_n, U1, U2, U3, blk, dt, mode, mask):
        phi_n = phi_n.contiguous()
        U = tuple(t.contiguous() for t in (U1, U2, U3))
        out = api.ad_RKTVD3(blk, phi_n, U, dt, adv_mask=mask, mode=mode)
        ctx.save_for_backward(phi_n, *U)
        ctx.blk, ctx.dt, ctx.mode, ctx.mask = blk, dt, mode, mask
        return out

    @staticmethod
    def backward(ctx, g):
        phi_n, U1, U2, U3 = ctx.saved_tensors
        want_u = any(ctx.needs_input_grad[1:4])
        aU = [torch.zeros_like(U1) for _ in range(3)] if want_u else None
        adj = api.ad_RKTVD3_adjoint(ctx.blk, phi_n, (U1, U2, U3), ctx.dt, g.contiguous(),
                                    adv_mask=ctx.mask, mode=ctx.mode, adj_ucn=aU)
        gu = aU if want_u else [None, None, None]
        return (adj, *gu, None, None, None, None)


def rk3_step(blk, phi_n, ucn: Sequence[torch.Tensor], dt: float, mode: str = "exact",
             adv_mask: Optional[torch.Tensor] = None):
    """One differentiable ad_RKTVD3 step."""
    return _RK3Step.apply(phi_n, ucn[0], ucn[1], ucn[2], blk, float(dt), mode, adv_mask)


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, U1, U2, U3, blk, dt, nsteps, mode, mask, every):
        U = tuple(t.contiguous() for t in (U1, U2, U3))
        ckpts = []
        cur = phi0.contiguous().clone()
        work = blk.workspace(api._lib.LSM_OP_RK3)
        for s in range(nsteps):
            if s % every == 0:
                ckpts.append(cur)
            nxt = api.ad_RKTVD3(blk, cur, U, dt, adv_mask=mask, mode=mode, work=work)
            cur = nxt
        ctx.save_for_backward(*U, *ckpts)
        ctx.blk, ctx.dt, ctx.nsteps, ctx.mode, ctx.mask, ctx.every = blk, dt, nsteps, mode, mask, every
        return cur

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        U, ckpts = saved[:3], saved[3:]
        blk, dt, nsteps, mode, mask, every = ctx.blk, ctx.dt, ctx.nsteps, ctx.mode, ctx.mask, ctx.every
        want_u = any(ctx.needs_input_grad[1:4])
        aU = [torch.zeros_like(U[0]) for _ in range(3)] if want_u else None
        work_f = blk.workspace(api._lib.LSM_OP_RK3)
        work_a = blk.workspace(api._lib.LSM_OP_RK3_ADJOINT)
        lam = g.contiguous().clone()
        nseg = len(ckpts)
        for sidx in range(nseg - 1, -1, -1):
            s0 = sidx * every
            s1 = min(s0 + every, nsteps)
            states = [ckpts[sidx]]
            for _ in range(s0, s1 - 1):
                states.append(api.ad_RKTVD3(blk, states[-1], U, dt, adv_mask=mask, mode=mode, work=work_f))
            for st in reversed(states):
                lam = api.ad_RKTVD3_adjoint(blk, st, U, dt, lam, adv_mask=mask, mode=mode, adj_ucn=aU,
                                            work=work_a)
        gu = aU if want_u else [None, None, None]
        return (lam, *gu, None, None, None, None, None, None)


def advect(blk, phi0, ucn: Sequence[torch.Tensor], dt: float, nsteps: int, mode: str = "exact",
           adv_mask: Optional[torch.Tensor] = None, checkpoint_every: int = 5):
    """nsteps differentiable ad_RKTVD3 steps with checkpointing every k steps."""
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    every = max(1, int(checkpoint_every))
    return _Trajectory.apply(phi0, ucn[0], ucn[1], ucn[2], blk, float(dt), int(nsteps), mode, adv_mask,
                             every)
