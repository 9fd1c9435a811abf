"""Many independent solitary-wave channels and the surface-elevation inverse
problem on them.

Each instance is the reference's channel configuration (inputs.channel_grid:
x and y uniform, z tanh-stretched toward the still-water depth d = 1; initial
phi = z - (d + H sech^2(k (x - x0))), frozen first-order solitary-wave
velocity) with its own grid size and wave amplitude H.  The instances advance
with one time step and one set of reinitialisation parameters (as one batch
run of the reference would be configured by one control file): the CFL step
of the fastest instance, dtau = 0.5 * the smallest spacing and epslsm = 1.5 *
the largest spacing over the batch, RNitermax iterations every step.

The loss is the squared misfit of the zero-level-set surface elevation
eta(x, y, t_end) = z_fs - d (z_fs from find_z_free_surface, the free-surface
search of calc_h) against each instance's own run from a perturbed amplitude;
its gradient with respect to each instance's initial phi and amplitude H
comes from the hand-written adjoints (levelset_trajectory with checkpoints,
free_surface's adjoint) and torch operations for phi0(H) and u(H).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np
import torch

from . import api, autograd, inputs


@dataclass
class ChannelSpec:
    nx: int
    ny: int
    nz: int
    H: float


def channel_specs(n: int = 32, seed: int = 8, nx=(128, 512), ny=(16, 64), nz=(48, 128),
                  H=(0.1, 0.5)) -> List[ChannelSpec]:
    """n instances with nx ~ U{nx}, ny ~ U{ny}, nz ~ U{nz}, H/d ~ U[H] from
    numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = int(rng.integers(nx[0], nx[1] + 1))
        b = int(rng.integers(ny[0], ny[1] + 1))
        c = int(rng.integers(nz[0], nz[1] + 1))
        out.append(ChannelSpec(a, b, c, float(rng.uniform(H[0], H[1]))))
    return out


def perturbed_amplitudes(specs: Sequence[ChannelSpec], seed: int = 9, rel: float = 0.1):
    """The target runs' amplitudes: H (1 + U[-rel, rel])."""
    rng = np.random.default_rng(seed)
    return [s.H * (1.0 + rng.uniform(-rel, rel)) for s in specs]


@dataclass
class ChannelBatch:
    specs: List[ChannelSpec]
    grids: list
    blocks: list
    dt: float
    params: api.ReinitParams
    depth: float = 1.0


def channel_batch(specs: Sequence[ChannelSpec], rn_itermax: int = 4, cfl: float = 0.4,
                  btype=(1, 1, 1, 1, 1, 1), big_phi: float = 5.0) -> ChannelBatch:
    """Grids, GPU blocks and the batch's common time step and reinitialisation
    parameters."""
    grids = [inputs.channel_grid(s.nx, s.ny, s.nz) for s in specs]
    dts, hmin, hmax = [], [], []
    for s, g in zip(specs, grids):
        u, v, w = inputs.solitary_velocity(g, H=s.H)
        dts.append(inputs.cfl_dt(g, u, v, w, cfl=cfl))
        lo, hi = inputs.spacing_bounds(g)
        hmin.append(lo)
        hmax.append(hi)
    prm = api.ReinitParams(epslsm=1.5 * max(hmax), RNtimestep=0.5 * min(hmin), RNitermax=rn_itermax,
                           epsReinitialisation=1e-12)
    blocks = [api.Block(g, btype=btype, big_phi=big_phi) for g in grids]
    return ChannelBatch(list(specs), grids, blocks, float(min(dts)), prm)


def instance_forward(cb: ChannelBatch, b: int, H, dphi=None, nsteps: int = 60, every: int = 10,
                     mode: str = "fast"):
    """phi at t_end of instance b from phi0 = solitary_phi(H) (+ dphi) under the
    solitary-wave velocity of amplitude H; differentiable in H and dphi."""
    blk = cb.blocks[b]
    phi0 = autograd.solitary_phi(blk, H, d=cb.depth)
    if dphi is not None:
        phi0 = phi0 + dphi
    U = autograd.contravariant(blk, *autograd.solitary_velocity(blk, H, d=cb.depth))
    return autograd.levelset_trajectory(blk, phi0, U, cb.dt, cb.params, nsteps, mode=mode,
                                        checkpoint_every=every)


def elevation(cb: ChannelBatch, b: int, phi):
    """eta = z_fs - d of every column of instance b, shape (ny, nx)."""
    return autograd.free_surface(cb.blocks[b], phi) - cb.depth


def targets(cb: ChannelBatch, H_target, nsteps: int = 60, every: int = 10, mode: str = "fast"):
    """eta(t_end) of each instance's run from its target amplitude."""
    out = []
    with torch.no_grad():
        for b in range(len(cb.blocks)):
            H = torch.tensor(float(H_target[b]), dtype=torch.float64, device=cb.blocks[b].device)
            out.append(elevation(cb, b, instance_forward(cb, b, H, None, nsteps, every, mode)))
    return out


def loss_and_gradients(cb: ChannelBatch, eta_t, nsteps: int = 60, every: int = 10, mode: str = "fast",
                       streams: int = 8):
    """J = sum_b mean_columns (eta_b - eta*_b)^2 and, per instance, dJ/dphi0
    (a field) and dJ/dH (total: through phi0(H) and u(H)).  The instances'
    forward and backward passes are issued round-robin on `streams` HIP streams
    (default priority), so small instances run side by side; the backward of
    each runs on the stream of its forward."""
    main = torch.cuda.current_stream()
    pool = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]
    for st in pool:
        st.wait_stream(main)
    Js, g_phi, Hs = [], [], []
    for b, blk in enumerate(cb.blocks):
        with torch.cuda.stream(pool[b % len(pool)]):
            H = torch.tensor(cb.specs[b].H, dtype=torch.float64, device=blk.device, requires_grad=True)
            dphi = blk.zeros().requires_grad_(True)
            eta = elevation(cb, b, instance_forward(cb, b, H, dphi, nsteps, every, mode))
            Jb = ((eta - eta_t[b]) ** 2).mean()
            Jb.backward()
            Js.append(Jb.detach())
            g_phi.append(dphi.grad)
            Hs.append(H)
    for st in pool:
        main.wait_stream(st)
    torch.cuda.synchronize()
    J = float(sum(float(j) for j in Js))
    return J, g_phi, [float(H.grad) for H in Hs]


def loss_and_gradients_batched(cb: ChannelBatch, eta_t, nsteps: int = 60, every: int = 10, mode: str = "fast"):
    """loss_and_gradients with all instances advanced together: the forward
    steps, the replayed steps and the reverse sweeps of the checkpointed
    trajectory each run as one launch per kernel over every instance
    (autograd.multi_trajectory); phi0(H), u(H) and the surface elevation stay
    per-instance torch operations."""
    Hs, dphis, phi0s, ucns = [], [], [], []
    for b, blk in enumerate(cb.blocks):
        H = torch.tensor(cb.specs[b].H, dtype=torch.float64, device=blk.device, requires_grad=True)
        dphi = blk.zeros().requires_grad_(True)
        Hs.append(H)
        dphis.append(dphi)
        phi0s.append(autograd.solitary_phi(blk, H, d=cb.depth) + dphi)
        ucns.append(autograd.contravariant(blk, *autograd.solitary_velocity(blk, H, d=cb.depth)))
    phiT = autograd.multi_trajectory(cb.blocks, phi0s, ucns, cb.dt, cb.params, nsteps, mode=mode,
                                     checkpoint_every=every)
    Js = [((elevation(cb, b, phiT[b]) - eta_t[b]) ** 2).mean() for b in range(len(cb.blocks))]
    torch.stack(Js).sum().backward()
    torch.cuda.synchronize()
    # the loss summed in instance order, as loss_and_gradients does
    return float(sum(float(j.detach()) for j in Js)), [d.grad for d in dphis], [float(H.grad) for H in Hs]
