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
* reinit(blk, phi0, params): reinitialisation_test2 (Sussman, ENO2) with the
  hand-written adjoint (lsm_reinit_sussman_eno2_adjoint).
* levelset_step(blk, phi_n, ucn, dt, params): one full step
  bcond_lsm(reinitialisation_test2(ad_RKTVD3(phi^n))) with its adjoint
  (lsm_levelset_step_adjoint); levelset_trajectory(...) runs nsteps of them
  with checkpointing every k steps; multi_trajectory(...) does the same for
  n independent blocks at once, one launch per kernel over all of them.

* free_surface(blk, phi): the free-surface height z_fs of every column
  (find_z_free_surface, as calc_h obtains it) with its adjoint
  (lsm_free_surface_adjoint); the surface elevation is z_fs - d.
* solitary_phi / solitary_velocity(blk, H): the solitary-wave channel's
  initial level set and frozen velocity in torch operations, differentiable
  in the amplitude H.

The adjoint differentiates the WENO3 nonlinear weights and s(phi0); the upwind
choice, the forced boundary stencils, the BigPhi switch of bcond_lsm, the ENO2
and one-sided branches of |grad d|, the reinitialisation update mask and its
iteration count are taken as in the forward (a.e. derivatives of the
piecewise-smooth step).
"""
from __future__ import annotations

import math
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


class _Reinit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, blk, params):
        phi0 = phi0.contiguous()
        out, _ = api.reinitialisation_test2(blk, phi0, params)
        ctx.save_for_backward(phi0)
        ctx.blk, ctx.params = blk, params
        return out

    @staticmethod
    def backward(ctx, g):
        (phi0,) = ctx.saved_tensors
        return api.reinitialisation_test2_adjoint(ctx.blk, phi0, ctx.params, g.contiguous()), None, None


def reinit(blk, phi0, params: "api.ReinitParams"):
    """Differentiable reinitialisation_test2 (Sussman ENO2, non-hybrid)."""
    return _Reinit.apply(phi0, blk, params)


def _step_forward(blk, phi_n, U, dt, params, mode, work_rk=None, work_rn=None, tmp=None):
    """bcond_lsm(reinitialisation_test2(ad_RKTVD3(phi_n)))."""
    X = api.ad_RKTVD3(blk, phi_n, U, dt, mode=mode, out=tmp, work=work_rk)
    out, _ = api.reinitialisation_test2(blk, X, params, work=work_rn)
    return api.bcond_lsm(blk, out)


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi_n, U1, U2, U3, blk, dt, params, mode):
        phi_n = phi_n.contiguous()
        U = tuple(t.contiguous() for t in (U1, U2, U3))
        out = _step_forward(blk, phi_n, U, dt, params, mode)
        ctx.save_for_backward(phi_n, *U)
        ctx.blk, ctx.dt, ctx.params, ctx.mode = blk, dt, params, mode
        return out

    @staticmethod
    def backward(ctx, g):
        phi_n, U1, U2, U3 = ctx.saved_tensors
        want_u = any(ctx.needs_input_grad[1:4])
        aU = [torch.zeros_like(U1) for _ in range(3)] if want_u else None
        adj = api.levelset_step_adjoint(ctx.blk, phi_n, (U1, U2, U3), ctx.dt, ctx.params, g.contiguous(),
                                        mode=ctx.mode, adj_ucn=aU)
        gu = aU if want_u else [None, None, None]
        return (adj, *gu, None, None, None, None)


def levelset_step(blk, phi_n, ucn: Sequence[torch.Tensor], dt: float, params: "api.ReinitParams",
                  mode: str = "exact"):
    """One differentiable level-set step: ad_RKTVD3, reinitialisation_test2,
    bcond_lsm (levelsetmethod with numiter_lsm = 1, reinitialisation every step)."""
    return _Step.apply(phi_n, ucn[0], ucn[1], ucn[2], blk, float(dt), params, mode)


class _StepTrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, U1, U2, U3, blk, dt, params, nsteps, mode, every, record):
        U = tuple(t.contiguous() for t in (U1, U2, U3))
        ckpts = []
        cur = phi0.contiguous().clone()
        wrk = blk.workspace(api._lib.LSM_OP_RK3)
        wrn = blk.workspace(api._lib.LSM_OP_REINIT_ENO2)
        tmp = blk.empty()
        # the steps of the last segment are recorded now (same results as the
        # plain steps), so the backward pass need not replay them
        last0 = ((nsteps - 1) // every) * every if record else nsteps
        rec_states, rec_traces = [], []
        for s in range(nsteps):
            if s % every == 0:
                ckpts.append(cur)
            if s >= last0:
                tr = blk.adjoint_workspace(api._lib.LSM_OP_STEP_TRACE, params.RNitermax)
                nxt = blk.empty()
                api.levelset_step_record(blk, cur, U, dt, params, mode=mode, phi_np1=nxt, trace=tr)
                rec_states.append(cur)
                rec_traces.append(tr)
                cur = nxt
            else:
                cur = _step_forward(blk, cur, U, dt, params, mode, wrk, wrn, tmp).clone()
        ctx.save_for_backward(*U, *ckpts)
        ctx.blk, ctx.dt, ctx.params, ctx.nsteps, ctx.mode, ctx.every = blk, dt, params, nsteps, mode, every
        ctx.rec = (rec_states, rec_traces)
        return cur

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        U, ckpts = saved[:3], saved[3:]
        blk, dt, params, nsteps, mode, every = ctx.blk, ctx.dt, ctx.params, ctx.nsteps, ctx.mode, ctx.every
        want_u = any(ctx.needs_input_grad[1:4])
        aU = [torch.zeros_like(U[0]) for _ in range(3)] if want_u else None
        wad = blk.adjoint_workspace(api._lib.LSM_OP_STEP_ADJOINT_RECORDED, params.RNitermax)
        lam = g.contiguous().clone()
        nxt = blk.empty()
        rec_states, traces = ctx.rec
        ctx.rec = None
        for sidx in range(len(ckpts) - 1, -1, -1):
            s0 = sidx * every
            s1 = min(s0 + every, nsteps)
            if sidx == len(ckpts) - 1 and rec_states:
                # recorded by the forward pass
                states = rec_states
            else:
                # replay the segment once, recording every step for its reverse sweep
                states = [ckpts[sidx]]
                for q in range(s1 - s0):
                    if len(traces) <= q:
                        traces.append(blk.adjoint_workspace(api._lib.LSM_OP_STEP_TRACE, params.RNitermax))
                    nxt_state = blk.empty() if q < s1 - s0 - 1 else None
                    api.levelset_step_record(blk, states[-1], U, dt, params, mode=mode, phi_np1=nxt_state,
                                             trace=traces[q])
                    if nxt_state is not None:
                        states.append(nxt_state)
            for q in range(s1 - s0 - 1, -1, -1):
                api.levelset_step_adjoint_recorded(blk, states[q], U, dt, params, traces[q], lam, mode=mode,
                                                   adj_ucn=aU, out=nxt, work=wad)
                lam, nxt = nxt, lam
        gu = aU if want_u else [None, None, None]
        return (lam, *gu, None, None, None, None, None, None, None)


def _wants_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts)


def levelset_trajectory(blk, phi0, ucn: Sequence[torch.Tensor], dt: float, params: "api.ReinitParams",
                        nsteps: int, mode: str = "exact", checkpoint_every: int = 5):
    """nsteps differentiable level-set steps (levelset_step) with checkpointing
    every k steps: the forward keeps nsteps/k checkpoints; the backward replays
    one segment at a time, recording each of its k steps (lsm_levelset_step_
    record) and then running their reverse sweeps from the records, so each
    step's forward is computed twice in all (forward pass, replay), except the
    last segment's, which the forward pass records when a gradient is wanted."""
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    every = max(1, int(checkpoint_every))
    return _StepTrajectory.apply(phi0, ucn[0], ucn[1], ucn[2], blk, float(dt), params, int(nsteps), mode,
                                 every, _wants_grad(phi0, *ucn))


class _MultiTrajectory(torch.autograd.Function):
    """levelset_trajectory of n independent blocks advanced together: every
    step (forward, replay with recording, reverse sweep) is one multi-instance
    execution (api.multi), i.e. one launch per kernel over all blocks."""

    @staticmethod
    def forward(ctx, blks, dt, params, nsteps, mode, every, record, *t):
        n = len(blks)
        phi0, U = t[:n], [tuple(t[n + 3 * b + c].contiguous() for c in range(3)) for b in range(n)]
        cur = [p.contiguous().clone() for p in phi0]
        nxt = [blk.empty() for blk in blks]
        tmp = [blk.empty() for blk in blks]
        wrk = [blk.workspace(api._lib.LSM_OP_RK3) for blk in blks]
        wrn = [blk.workspace(api._lib.LSM_OP_REINIT_ENO2) for blk in blks]
        ckpts = []
        # the steps of the last segment are recorded now (same results as the
        # plain steps), so the backward pass need not replay them
        last0 = ((nsteps - 1) // every) * every if record else nsteps
        rec_states, rec_traces = [], []
        for s in range(nsteps):
            if s % every == 0:
                ckpts.append([c.clone() for c in cur])
            if s >= last0:
                trs = [blk.adjoint_workspace(api._lib.LSM_OP_STEP_TRACE, params.RNitermax) for blk in blks]
                new = [blk.empty() for blk in blks]
                with api.multi() as m:
                    for b, blk in enumerate(blks):
                        m.instance()
                        api.levelset_step_record(blk, cur[b], U[b], dt, params, mode=mode, phi_np1=new[b],
                                                 trace=trs[b])
                rec_states.append(cur)
                rec_traces.append(trs)
                cur = new
                continue
            with api.multi() as m:
                for b, blk in enumerate(blks):
                    m.instance()
                    x = api.ad_RKTVD3(blk, cur[b], U[b], dt, mode=mode, out=tmp[b], work=wrk[b])
                    api.reinitialisation_test2(blk, x, params, out=nxt[b], work=wrn[b])
                    api.bcond_lsm(blk, nxt[b])
            cur, nxt = nxt, cur
        flat_u = [u for b in range(n) for u in U[b]]
        ctx.save_for_backward(*flat_u, *[c for seg in ckpts for c in seg])
        ctx.blks, ctx.dt, ctx.params, ctx.nsteps, ctx.mode, ctx.every = blks, dt, params, nsteps, mode, every
        ctx.nseg = len(ckpts)
        ctx.rec = (rec_states, rec_traces)
        return tuple(cur)

    @staticmethod
    def backward(ctx, *g):
        blks, dt, params, nsteps, mode, every = ctx.blks, ctx.dt, ctx.params, ctx.nsteps, ctx.mode, ctx.every
        n = len(blks)
        saved = ctx.saved_tensors
        U = [tuple(saved[3 * b + c] for c in range(3)) for b in range(n)]
        ck = saved[3 * n:]
        ckpts = [ck[q * n:(q + 1) * n] for q in range(ctx.nseg)]
        want_u = any(ctx.needs_input_grad[7 + n:])
        aU = [[torch.zeros_like(U[b][0]) for _ in range(3)] for b in range(n)] if want_u else [None] * n
        wad = [blk.adjoint_workspace(api._lib.LSM_OP_STEP_ADJOINT_RECORDED, params.RNitermax) for blk in blks]
        lam = [(gb if gb is not None else torch.zeros_like(U[b][0])).contiguous().clone() for b, gb in enumerate(g)]
        nxt = [blk.empty() for blk in blks]
        rec_states, traces = ctx.rec
        ctx.rec = None
        for sidx in range(ctx.nseg - 1, -1, -1):
            s0 = sidx * every
            s1 = min(s0 + every, nsteps)
            # replay the segment once for all blocks, recording every step
            # (the last segment was recorded by the forward pass)
            states = [list(ckpts[sidx])]
            if sidx == ctx.nseg - 1 and rec_states:
                states = rec_states
            for q in range(s1 - s0 if states is not rec_states else 0):
                if len(traces) <= q:
                    traces.append([blk.adjoint_workspace(api._lib.LSM_OP_STEP_TRACE, params.RNitermax)
                                   for blk in blks])
                last = q == s1 - s0 - 1
                nstate = None if last else [blk.empty() for blk in blks]
                with api.multi() as m:
                    for b, blk in enumerate(blks):
                        m.instance()
                        api.levelset_step_record(blk, states[-1][b], U[b], dt, params, mode=mode,
                                                 phi_np1=None if last else nstate[b], trace=traces[q][b])
                if not last:
                    states.append(nstate)
            for q in range(s1 - s0 - 1, -1, -1):
                with api.multi() as m:
                    for b, blk in enumerate(blks):
                        m.instance()
                        api.levelset_step_adjoint_recorded(blk, states[q][b], U[b], dt, params, traces[q][b],
                                                           lam[b], mode=mode, adj_ucn=aU[b], out=nxt[b],
                                                           work=wad[b])
                lam, nxt = nxt, lam
        gu = [aU[b][c] if want_u else None for b in range(n) for c in range(3)]
        return (None, None, None, None, None, None, None, *lam, *gu)


def multi_trajectory(blks, phi0s, ucns, dt: float, params: "api.ReinitParams", nsteps: int,
                     mode: str = "exact", checkpoint_every: int = 5):
    """levelset_trajectory for n independent blocks (phi0s[b], ucns[b] = (U1,
    U2, U3) of block b), advanced together: each step, replayed step and
    reverse sweep runs as one launch per kernel over all blocks.  Returns the
    n final fields."""
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    flat = [u for U in ucns for u in U]
    return _MultiTrajectory.apply(list(blks), float(dt), params, int(nsteps), mode,
                                  max(1, int(checkpoint_every)), _wants_grad(*phi0s, *flat), *phi0s, *flat)


def heaviside(phi, eps: float):
    """Smoothed Heaviside H_eps (global_lsm.F90:105-118)."""
    mid = 0.5 * (1.0 + phi / eps + (1.0 / math.pi) * torch.sin(math.pi * phi / eps))
    return torch.where(phi < -eps, torch.zeros_like(phi), torch.where(phi > eps, torch.ones_like(phi), mid))


def fourier_velocity(blk, ks, es, amps, psis):
    """u = sum_m amps_m e_m cos(2 pi k_m . x + psi_m) at the block's nodes, in
    torch operations (amps may require grad)."""
    x, y, z = blk.x, blk.y, blk.z
    u = torch.zeros_like(x)
    v = torch.zeros_like(x)
    w = torch.zeros_like(x)
    for m in range(len(ks)):
        k, e = ks[m], es[m]
        ph = torch.cos(2.0 * math.pi * (float(k[0]) * x + float(k[1]) * y + float(k[2]) * z) + float(psis[m]))
        u = u + amps[m] * float(e[0]) * ph
        v = v + amps[m] * float(e[1]) * ph
        w = w + amps[m] * float(e[2]) * ph
    return u, v, w


class _FreeSurface(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, blk, z):
        zfs, _ = api.free_surface(blk, z, phi.detach().contiguous())
        ctx.blk, ctx.z = blk, z
        ctx.save_for_backward(phi.detach().contiguous())
        return zfs

    @staticmethod
    def backward(ctx, g):
        (phi,) = ctx.saved_tensors
        return api.free_surface_adjoint(ctx.blk, ctx.z, phi, g.contiguous()), None, None


def free_surface(blk, phi, z: Optional[torch.Tensor] = None):
    """z_fs of every (i, j) column, shape (nj, ni): the lowest sign change of
    phi in z, linearly interpolated (find_z_free_surface.F90:16-58; 0 in a
    column without one).  Differentiable in phi: only the interpolating case
    depends on phi, along the interval the forward found."""
    return _FreeSurface.apply(phi, blk, blk.z if z is None else z)


def _solitary(blk, H, d, x0):
    """sech^2 profile of the first-order solitary wave (the formulas of
    inputs.solitary_wave) in torch operations."""
    k = torch.sqrt(3.0 * H / (4.0 * d ** 3))
    arg = k * (blk.x - x0)
    s = 1.0 / torch.cosh(arg)
    eta = H * s * s
    deta = -2.0 * H * k * s * s * torch.tanh(arg)
    return eta, deta


def solitary_phi(blk, H, d: float = 1.0, x0: float = 5.0):
    """phi = z - (d + H sech^2(k (x - x0))), k = sqrt(3 H / (4 d^3))."""
    eta, _ = _solitary(blk, H, d, x0)
    return blk.z - (d + eta)


def solitary_velocity(blk, H, d: float = 1.0, x0: float = 5.0, gacc: float = 9.81):
    """u = c eta / d, v = 0, w = -(c / d) eta'(x) z, c = sqrt(g (d + H))."""
    c = torch.sqrt(gacc * (d + H))
    eta, deta = _solitary(blk, H, d, x0)
    u = c * eta / d
    w = -(c / d) * deta * blk.z
    return u, torch.zeros_like(u), w
