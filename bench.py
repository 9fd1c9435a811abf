#!/usr/bin/env python3
"""Level-set transport benchmark on one MI355X.

Workload (the reference's 256^3 forward case): one 256^3 curvilinear block on
the unit cube (x_a = xi_a + 0.03 sin(2 pi xi_b + t1) sin(2 pi xi_c + t2),
seeded phases), phi0 = signed distance to 8 random spheres, a frozen
divergence-free random Fourier velocity (16 modes, |k| <= 4, unit RMS), SSP-RK3
WENO3 steps at CFL 0.4, FP64.  A step is one ad_RKTVD3 call: three fused
stage kernels plus three boundary-shell kernels.

Prints one JSON line (metric, value = advection cell-updates/s, roofline of
the stage kernel, CPU baseline of the reference on the host).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ns-lsm-single-phase_amd"))

METRIC = "level-set cell-updates/s (fwd, adjoint) FP64 at 256³; % HBM roofline"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per interior cell and RK stage (the reference data model):
# stage 1 reads phi^n, U/J (3), aj and writes phi (48 B); stages 2-3 also read
# the stage input phi (56 B): 160 B per advection cell-update
ADV_BYTES_PER_CELL_UPDATE = 160.0
ADV_FOLDED_BYTES_PER_CELL_UPDATE = 136.0
VOLUME_BYTES_PER_CELL = 32.0
# Sussman iteration on the metric-tensor data model the kernels use: stage 1
# reads d, phi0 and the six tensor components and writes d and the update
# mask (73 B); stages 2-3 also read d^n and the mask (81 B).  (The reference
# data model, nine metric arrays per node, is 304 B.)
REINIT_BYTES_PER_CELL_ITER = 235.0
# algorithmic bytes per node of the reverse-sweep kernels, averaged over the
# three RK stages of a launch sequence.  Advection adjoint stage: reads the
# stage input, the adjoint, U/J (3), aj, the accumulated dJ/d(U/J) (3) and
# dJ/dphi^n; writes dJ/d(U/J) (3), dJ/dphi^n and (stages 3, 2) the adjoint of
# the stage input: 120, 120, 112 B.  Tape gather: reads the adjoint, the three
# tape weights, the branch code (2 B) and the update mask (1 B), writes the
# adjoint of the stage input, and the iteration's dJ/dd^n (written, updated,
# read over stages 3, 2, 1: 8, 16, 8 B).
ADJ_STAGE_BYTES_PER_NODE = (120.0 + 120.0 + 112.0) / 3.0
TAPE_GATHER_BYTES_PER_NODE = 8 + 24 + 2 + 1 + 8 + (8.0 + 16.0 + 8.0) / 3.0



def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(n, seed_grid=2, seed_phi=3, seed_vel=4):
    import numpy as np
    from lsm_amd import inputs
    g = inputs.sinusoidal_cube_grid(n, amp=0.03, seed=seed_grid)
    phi = inputs.spheres_sdf(g, nspheres=8, seed=seed_phi)
    vel = inputs.fourier_velocity(g, nmodes=16, kmax=4, seed=seed_vel)
    dt = inputs.cfl_dt(g, *vel, cfl=0.4)
    return g, np.ascontiguousarray(phi), vel, dt


def cpu_baseline(n_ij=256, nk=64, steps=6):
    """The reference Fortran (oracle/_ref, compiled with amdflang) on the
    host cores: one MPI rank per core, every rank advancing its own copy of a
    slab of the same i-j extent (the single-block case each rank of a domain
    decomposition owns, without the halo exchange), so the aggregate rate is
    an upper bound for the reference on these cores.  Falls back to the C
    restatement of the oracle on one core if the Fortran build is absent."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from lsm_amd import inputs
    g = inputs.sinusoidal_cube_grid(n_ij, seed=2, nk=nk)
    phi = inputs.spheres_sdf(g, seed=3)
    u, v, w = inputs.fourier_velocity(g, seed=4)
    dt = inputs.cfl_dt(g, u, v, w)
    cells = (n_ij - 2) * (n_ij - 2) * (nk - 2) * steps
    ranks = max(1, min(16, os.cpu_count() or 1))
    sample = (f"{steps} ad_RKTVD3 steps on a {n_ij}x{n_ij}x{nk} block of the same mapping "
              f"({cells} cell-updates per rank)")
    try:
        import refio
        if refio.available():
            _, secs, _ = refio.run("advect", g, phi, u=u, v=v, w=w, delti_lsm=dt, nsteps=steps,
                                   timing=True, nprocs=ranks)
            if secs:
                return {"value": ranks * cells / secs, "unit": "cell-updates/s", "cores": ranks,
                        "kind": "reference", "seconds": secs,
                        "sample": sample + f", compiled reference Fortran (amdflang -O3, the reference's own flag), {ranks} MPI "
                                           "ranks, one per core, each on its own block; slowest rank timed"}
    except Exception as e:  # noqa: BLE001
        log(f"reference harness unavailable ({e}); timing the C oracle instead")
    import oracle as O
    blk = O.Block(g)
    U = O.contra(blk, u, v, w)
    t0 = time.perf_counter()
    O.rk3(blk, phi, U, dt, nsteps=steps)
    secs = time.perf_counter() - t0
    return {"value": cells / secs, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "seconds": secs, "sample": sample + ", C restatement of the oracle (gcc -O2, 1 thread)"}


def cpu_geometric_baseline():
    """The reference's hybrid (geometric) reinitialisation, compiled Fortran
    (oracle/_ref), on the host cores: a levelsetmethod call with
    call_levelsetmethod = .false. (water volume, GetTetrahedraList,
    GetIsosurfaceAndDistances, NeighboursPhiCorrection, bcond) on a 32 x 32 x 24
    block of the bench's mapping and level set, one MPI rank per core, each on
    its own copy; the slowest rank is timed.  The band nodes (interface nodes
    and their neighbours) of the sample come from the GPU path on the same
    input.  None when the reference build is absent."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    from lsm_amd import api, inputs
    try:
        import refio
        if not refio.available():
            return None
    except Exception:  # noqa: BLE001
        return None
    g = inputs.sinusoidal_cube_grid(32, seed=2, nk=24)
    phi = inputs.spheres_sdf(g, seed=3)
    z = np.zeros(g.shape)
    ranks = max(1, min(16, os.cpu_count() or 1))
    _, secs, _ = refio.run("levelset", g, phi, phi0=phi, u=z, v=z, w=z, hybrid=True, call_levelset=False,
                           rn_freq=1, timing=True, epslsm=0.05, nprocs=ranks)
    if not secs:
        return None
    blk = api.Block(g)
    p = torch.as_tensor(phi, device="cuda")
    pc = torch.where(p < -2.220446049250313e-16, torch.full_like(p, -9.9),
                     torch.where(p > 2.220446049250313e-16, torch.full_like(p, 9.9), torch.zeros_like(p)))
    ids = blk.empty(torch.int32)
    api.geometric_reinitialisation(blk, p, pc, ids)
    band = int((ids != 0).sum())
    return {"value": ranks * band / secs, "unit": "band-nodes/s", "cores": ranks, "kind": "reference",
            "seconds": secs, "sample": f"hybrid levelsetmethod without advection on a 32x32x24 block "
                                       f"({band} band nodes per rank), {ranks} MPI ranks, one per core"}


def bench_fortran_route(n=128, steps=5):
    """levelsetmethod called from Fortran on host arrays (the reference
    harness, one rank, levelsetmethod.F90 called as mg_driver.F90:563-581 does)
    at n^3 of the bench construction, 4 Sussman iterations per step: the
    unmodified levelsetmethod.F90 with the callee drop-ins (phi crosses PCIe
    at every callee), the whole-call drop-in (phi, phi_n and q up once, phi
    down once per call) and the compiled reference on one core; next to the
    same step on device-resident tensors.  Harness times include the first
    call's grid upload."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    from lsm_amd import api, inputs, _lib
    import refio
    g, phi, vel, dt = make_case(n)
    hmin, hmax = inputs.spacing_bounds(g)
    kw = dict(u=vel[0], v=vel[1], w=vel[2], delti_lsm=dt, nsteps=steps, epslsm=1.5 * hmax,
              rn_timestep=0.5 * hmin, eps_reinit=1e-12, rn_itermax=4, timing=True, phi0=phi)
    res = {"workload": f"{n}^3 of the bench construction, {steps} levelsetmethod calls (RK3 + 4 Sussman "
                       "iterations + volume + bcond), host arrays, one rank"}
    for tag, h in (("callee_dropins", refio.SHIM_HARNESS), ("whole_call_dropin", refio.RESIDENT_HARNESS),
                   ("reference_1core", refio.HARNESS)):
        if not os.path.exists(h):
            continue
        _, secs, _ = refio.run("levelset", g, phi, harness=h, **kw)
        if secs:
            res[f"{tag}_ms_per_call"] = round(secs * 1e3 / steps, 3)
    blk = api.Block(g)
    prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4, epsReinitialisation=1e-12)
    uvw = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in vel]
    p = torch.as_tensor(phi, device="cuda")
    pn = p.clone()
    lw = blk.workspace(_lib.LSM_OP_LEVELSET_STEP)
    res["device_resident_ms_per_call"] = round(_time(lambda: api.levelsetmethod(blk, p, pn, *uvw, prm, dt,
                                                                                 work=lw), steps), 3)
    return res


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _time(fn, reps, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_geometric(g, phi, blk, vel, dt):
    """The geometric (hybrid) reinitialisation at the bench size:
    geometric_reinitialisation alone (cut-cell search, isosurfaces, interface
    and neighbour flags, distances; one stream synchronisation) and a whole
    hybrid levelsetmethod call (AdvectionNodes band, band-restricted RK3,
    geometric reinitialisation, bcond).  The distance kernel is bound by FP64
    arithmetic and L2 traffic, not HBM: its figure is band nodes per second."""
    import ctypes as C
    import numpy as np
    import torch
    from lsm_amd import api, _lib
    L = _lib.load()
    pc0 = torch.where(phi < -2.220446049250313e-16, torch.full_like(phi, -9.9),
                      torch.where(phi > 2.220446049250313e-16, torch.full_like(phi, 9.9), torch.zeros_like(phi)))
    pc = pc0.clone()
    ids = blk.empty(torch.int32)
    work = blk.workspace(_lib.LSM_OP_REINIT_GEOMETRIC)
    api.geometric_reinitialisation(blk, phi, pc, ids, work=work)
    torch.cuda.synchronize()
    nint = int((ids == 1).sum())
    nnb = int((ids == -1).sum())
    L.lsm_timing_enable(1)

    def one():
        pc.copy_(pc0)
        api.geometric_reinitialisation(blk, phi, pc, None, work=work)
    ms = _time(one, 5)
    kms, kn = C.c_double(0), C.c_int64(0)
    L.lsm_timing_get(_lib.LSM_KT_GEOMETRIC, C.byref(kms), C.byref(kn))
    L.lsm_timing_enable(0)
    prm = api.ReinitParams(epslsm=0.05, RNtimestep=1e-3, RNitermax=4, epsReinitialisation=1e-5,
                           hybrid_reinitialisation=True)
    uvw = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in vel]
    pl = phi.clone()
    pn = phi.clone()
    lw = blk.workspace(_lib.LSM_OP_LEVELSET_STEP)

    def step():
        pn.copy_(phi)
        api.levelsetmethod(blk, pl, pn, *uvw, prm, dt, ntime=1, want_volume=False, work=lw)
    ls_ms = _time(step, 3)
    dist_ms = kms.value / max(kn.value, 1)
    return {"interface_nodes": nint, "neighbour_nodes": nnb, "reinit_ms": round(ms, 3),
            "distance_kernel_ms": round(dist_ms, 3),
            "band_nodes_per_s": (nint + nnb) / (ms / 1e3),
            "hybrid_levelsetmethod_ms": round(ls_ms, 3)}


def bench_step_adjoint(g, phi, ucn, dt, blk):
    """One full level-set step (ad_RKTVD3, 4 Sussman iterations, bcond_lsm)
    and its hand-written adjoint (lsm_levelset_step_adjoint: forward
    recompute with the reinitialisation trace, then the reverse sweep) at the
    bench size."""
    from lsm_amd import api, inputs, _lib
    hmin, hmax = inputs.spacing_bounds(g)
    prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4, epsReinitialisation=1e-12)
    wrk, wrn = blk.workspace(_lib.LSM_OP_RK3), blk.workspace(_lib.LSM_OP_REINIT_ENO2)
    wad = blk.adjoint_workspace(_lib.LSM_OP_STEP_ADJOINT, 4)
    tmp, out, adj = blk.empty(), blk.empty(), blk.empty()
    lam = phi.clone()
    aU = [blk.zeros() for _ in range(3)]

    def fwd():
        x = api.ad_RKTVD3(blk, phi, ucn, dt, mode="fast", out=tmp, work=wrk)
        api.reinitialisation_test2(blk, x, prm, out=out, work=wrn)
        api.bcond_lsm(blk, out)

    def bwd():
        api.levelset_step_adjoint(blk, phi, ucn, dt, prm, lam, mode="fast", adj_ucn=aU, out=adj, work=wad)
    # forward and adjoint timed in alternating rounds (the median of each), so
    # that both see the same clock and thermal state
    f_runs, b_runs = [], []
    for _ in range(3):
        f_runs.append(_time(fwd, 5))
        b_runs.append(_time(bwd, 3))
    f_ms, b_ms = sorted(f_runs)[1], sorted(b_runs)[1]
    n_int = (g.ni - 2) * (g.nj - 2) * (g.nk - 2)
    # the reverse-sweep kernels' own durations (library HIP events, one more call)
    import ctypes as C
    L = _lib.load()
    L.lsm_timing_enable(1)
    bwd()
    import torch
    torch.cuda.synchronize()
    kt = {}
    for name, kid in (("adj_stage", _lib.LSM_KT_ADJOINT_STAGE), ("tape_gather", _lib.LSM_KT_REINIT_ADJOINT)):
        ms, cnt = C.c_double(), C.c_int64()
        L.lsm_timing_get(kid, C.byref(ms), C.byref(cnt))
        kt[name] = (ms.value / max(cnt.value, 1), int(cnt.value))
    L.lsm_timing_enable(0)
    nodes = g.ni * g.nj * g.nk
    adj_roof = {}
    for name, per_node, prof in (("adj_stage", ADJ_STAGE_BYTES_PER_NODE, "adj_stage_lean_256"),
                                 ("tape_gather", TAPE_GATHER_BYTES_PER_NODE, "radj_tape_gather_256")):
        ms, cnt = kt[name]
        t = load_traffic(prof)
        traffic = t["bytes_per_launch"] if t and nodes == 256 ** 3 else None
        if cnt:
            alg = per_node * nodes
            gbs = alg / (ms / 1e3) / 1e9
            adj_roof[name] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_launch_ms": round(ms, 5), "launches": cnt,
                              "alg_bytes_per_launch": alg, "alg_bytes_per_node": per_node,
                              "traffic": traffic}
    return {"step_fwd_ms": round(f_ms, 4), "step_adjoint_ms": round(b_ms, 4),
            "step_adjoint_forward_ratio": round(b_ms / f_ms, 3),
            "adjoint_roofline": adj_roof,
            "step_adjoint_cell_updates_per_s": n_int / (b_ms / 1e3),
            "step_definition": "ad_RKTVD3 + 4 reinitialisation_test2 iterations + bcond_lsm (levelsetmethod, "
                               "numiter_lsm = 1); adjoint = forward recompute incl. the reinitialisation trace + "
                               "reverse sweep, gradients w.r.t. phi^n and U/J"}


def bench_consumers(g, phi, blk):
    """The caller-side consumers of phi at the bench size: calc_phi_gradient_ENO2
    (phi + 9 metrics in, 3 components out: 104 B per node), calc_h (phi and z
    in, h out: 24 B per node) and calc_h_gradient_ENO2 (h + 6 metrics in, 2
    components out: 72 B per node)."""
    import numpy as np
    import torch
    from lsm_amd import api
    n = g.ni * g.nj * g.nk
    z = torch.as_tensor(np.ascontiguousarray(g.z), device="cuda")
    gp = [blk.zeros() for _ in range(3)]
    gh = [blk.zeros() for _ in range(2)]
    h = blk.empty()
    res = {}
    for name, fn, nbytes in (
            ("phi_gradient", lambda: api.calc_phi_gradient_ENO2(blk, phi, 2, True, out=gp), 104),
            ("calc_h", lambda: api.calc_h(blk, z, phi, out=h), 24),
            ("h_gradient", lambda: api.calc_h_gradient_ENO2(blk, h, 2, True, out=gh), 72)):
        ms = _time(fn, 10)
        res[f"{name}_ms"] = round(ms, 4)
        res[f"{name}_alg_GBps"] = round(n * nbytes / (ms / 1e3) / 1e9, 1)
    return res


def bench_batch(steps=10):
    """The 64-block workload: a 4x4x4 lattice of curvilinear blocks, sizes
    U{32..160} per lattice slab (seed 6), per-block sinusoidal perturbation
    amplitudes U[0, 0.04], 3 ghost layers, 32 spheres, divergence-free Fourier
    velocity; per-block plain launches against the batched form (autotuned
    stage tiles, hipGraph per step), advection-only and full steps (3 RK
    stages + 4 Sussman iterations + bcond + halo fill)."""
    import torch
    from lsm_amd import api, batch as B, inputs
    t0 = time.time()
    lay, grids, phi, vel, _ = B.multiblock_case()
    dt = B.cfl_dt(lay, grids, vel)
    hmin = min(inputs.spacing_bounds(gd)[0] for gd in grids)
    hmax = max(inputs.spacing_bounds(gd)[1] for gd in grids)
    prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4, epsReinitialisation=1e-12)
    res = {"workload": "64 blocks (4x4x4 lattice, sizes U{32..160}, 3 ghost layers), 32 spheres, "
                       "16-mode Fourier velocity, FP64", "cells": lay.cells(), "sizes": lay.sizes}
    # per-block plain launches on per-block arrays with dense rows (each block
    # as a rank of the reference holds it), one launch per block and kernel
    dlay = B.lattice_layout(lay.sizes, ghost=lay.ghost, align=0)
    db = B.Batch(dlay, grids)
    ducn = db.contra_velocity(*[db.pack([v[c] for v in vel]) for c in range(3)])
    dp0 = db.halo_fill(db.pack(phi))
    dout = db.empty()
    dw_adv, dw_full = db.workspace(), db.workspace(api._lib.LSM_OP_LEVELSET_STEP)
    db.set_config("perblock")
    res["perblock_adv_ms"] = round(_time(lambda: db.levelset_step(dp0, ducn, dt, out=dout, work=dw_adv), steps), 4)
    # the per-block form at its best reinitialisation tile width
    pb = {}
    for w in (64, 32, 16):
        db.set_reinit_tile(w)
        pb[w] = round(_time(lambda: db.levelset_step(dp0, ducn, dt, prm, out=dout, work=dw_full), steps), 4)
    res["perblock_full_ms_by_reinit_tile"] = pb
    res["perblock_full_ms"] = min(pb.values())
    del db, ducn, dp0, dout, dw_adv, dw_full
    torch.cuda.empty_cache()
    # the batch: rows padded to 128-byte lines, first owned column aligned
    bt = B.Batch(lay, grids)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    p0 = bt.halo_fill(bt.pack(phi))
    del grids, vel, phi
    log(f"batch inputs: {len(lay.blocks)} blocks, {lay.cells()} cells, {time.time() - t0:.1f}s")
    out = bt.empty()
    w_adv, w_full = bt.workspace(), bt.workspace(api._lib.LSM_OP_LEVELSET_STEP)
    bt.set_config("perblock")
    res["perblock_aligned_adv_ms"] = round(_time(lambda: bt.levelset_step(p0, ucn, dt, out=out, work=w_adv), steps), 4)
    pb = {}
    for w in (64, 32, 16):
        bt.set_reinit_tile(w)
        pb[w] = round(_time(lambda: bt.levelset_step(p0, ucn, dt, prm, out=out, work=w_full), steps), 4)
    res["perblock_aligned_full_ms"] = min(pb.values())
    bt.set_reinit_tile(0)
    cfg, ms = bt.autotune(p0, ucn, dt, reps=3, force=True, out=out, work=w_adv)
    res["autotuned"] = cfg
    best_w, times = bt.autotune_reinit(p0, ucn, dt, prm, reps=steps, out=out, work=w_full)
    res["full_ms_by_reinit_tile"] = {w: round(t, 4) for w, t in times.items()}
    res["reinit_tile"] = best_w
    res["batched_adv_ms"] = round(_time(lambda: bt.levelset_step(p0, ucn, dt, out=out, work=w_adv), steps), 4)
    res["batched_full_ms"] = round(_time(lambda: bt.levelset_step(p0, ucn, dt, prm, out=out, work=w_full), steps), 4)
    bt.graph_capture(out, p0, ucn, dt, work=w_adv)
    res["graph_adv_ms"] = round(_time(bt.graph_launch, steps), 4)
    bt.graph_capture(out, p0, ucn, dt, params=prm, work=w_full)
    res["graph_full_ms"] = round(_time(bt.graph_launch, steps), 4)
    for kind in ("adv", "full"):
        best = min(res[f"batched_{kind}_ms"], res[f"graph_{kind}_ms"])
        res[f"{kind}_cell_updates_per_s_batched"] = lay.cells() / (best / 1e3)
        res[f"{kind}_cell_updates_per_s_perblock"] = lay.cells() / (res[f"perblock_{kind}_ms"] / 1e3)
        res[f"{kind}_speedup_vs_perblock"] = round(res[f"perblock_{kind}_ms"] / best, 3)
        res[f"{kind}_speedup_vs_perblock_aligned"] = round(res[f"perblock_aligned_{kind}_ms"] / best, 3)
    bt.set_config("queue", 8, 8, 0)
    res["queue_adv_ms"] = round(_time(lambda: bt.levelset_step(p0, ucn, dt, out=out, work=w_adv), steps), 4)
    log(f"batch: adv {res['perblock_adv_ms']} -> {res['graph_adv_ms']} ms ({res['adv_speedup_vs_perblock']}x), "
        f"full {res['perblock_full_ms']} -> {res['graph_full_ms']} ms ({res['full_speedup_vs_perblock']}x)")
    del bt, ucn, p0, out, w_adv, w_full
    return res


def bench_inverse(nsteps=60, every=10, ninst=32):
    """The batched inverse problem: 32 solitary-wave channels (nx ~ U{128..512},
    ny ~ U{16..64}, nz ~ U{48..128}, H ~ U[0.1, 0.5], seed 8) with one time
    step (the CFL 0.4 step of the fastest instance) and 4 reinitialisation
    iterations per step; loss = sum over instances of the mean squared misfit
    of the surface elevation eta(x, y, t_end) against the instance's run from a
    perturbed amplitude (seed 9); gradients with respect to each instance's
    initial phi and amplitude H, checkpoint every 10 steps.  The forward runs
    as one independent batch (one launch per stage over all instances); the
    gradient's forward, replay and reverse sweeps run as multi-instance
    executions (one launch per kernel over all instances), and the same
    gradient issued instance by instance on 8 streams is timed beside it."""
    import torch
    from lsm_amd import _lib, batch as B, channels, inputs
    t0 = time.time()
    specs = channels.channel_specs(ninst, seed=8)
    cb = channels.channel_batch(specs)
    cells = sum(s.nx * s.ny * s.nz for s in specs)
    lay = B.independent_layout([(s.nx, s.ny, s.nz) for s in specs])
    bt = B.Batch(lay, cb.grids)
    vel = [inputs.solitary_velocity(g, H=s.H) for g, s in zip(cb.grids, specs)]
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    p0 = bt.pack([inputs.solitary_phi(g, H=s.H) for g, s in zip(cb.grids, specs)])
    del vel
    log(f"inverse inputs: {ninst} channels, {cells} cells, {time.time() - t0:.1f}s")
    bufs = [p0.clone(), bt.empty()]
    work = bt.workspace(_lib.LSM_OP_LEVELSET_STEP)

    def forward():
        for n in range(nsteps):
            bt.levelset_step(bufs[n % 2], ucn, cb.dt, cb.params, out=bufs[(n + 1) % 2], work=work)
    fwd_ms = _time(forward, 1, warm=1)
    del bufs, work, bt, ucn, p0
    torch.cuda.empty_cache()
    Ht = channels.perturbed_amplitudes(specs, seed=9)
    eta_t = channels.targets(cb, Ht, nsteps, every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J, g_phi, g_H = channels.loss_and_gradients_batched(cb, eta_t, nsteps, every)
    torch.cuda.synchronize()
    grad_ms = (time.perf_counter() - t1) * 1e3
    del g_phi
    torch.cuda.empty_cache()
    # the same gradient with the instances issued one by one on 8 streams
    t1 = time.perf_counter()
    J_s, _, g_H_s = channels.loss_and_gradients(cb, eta_t, nsteps, every)
    torch.cuda.synchronize()
    streams_ms = (time.perf_counter() - t1) * 1e3
    same = J_s == J and g_H_s == g_H
    return {"workload": f"{ninst} solitary-wave channels (nx U{{128..512}}, ny U{{16..64}}, nz U{{48..128}}, "
                        f"H U[0.1,0.5]), {nsteps} steps x (RK3 + 4 reinitialisation iterations + bcond), "
                        f"surface-elevation misfit, gradients w.r.t. phi0 and H, checkpoint every {every}",
            "cells": cells, "dt": cb.dt,
            "batched_forward_ms": round(fwd_ms, 2),
            "batched_forward_cell_updates_per_s": cells * nsteps / (fwd_ms / 1e3),
            "loss": J, "dJ_dH_first": g_H[:4],
            "forward_plus_gradient_ms": round(grad_ms, 1),
            "gradient_over_batched_forward": round(grad_ms / fwd_ms, 3),
            "gradient_form": "multi-instance: one launch per kernel over all instances, forward, replay and "
                             "reverse sweep",
            "per_instance_streams_forward_plus_gradient_ms": round(streams_ms, 1),
            "per_instance_streams_same_result": bool(same)}


def bench_differentiation(n=192, nsteps=20, every=5):
    """The differentiation workload at n^3: nsteps full steps (RK3 + 4 Sussman
    iterations + bcond), checkpoint every `every` steps, loss J = sum_cells
    V (H_eps(phi_T) - H_eps(phi_target))^2 (V = 1/aj, eps = 1.5 max spacing,
    phi_target: the same run from perturbed sphere radii), gradients with
    respect to phi0 and the 16 velocity Fourier amplitudes."""
    import numpy as np
    import torch
    from lsm_amd import api, autograd, inputs
    t0 = time.time()
    g = inputs.sinusoidal_cube_grid(n, amp=0.03, seed=2)
    phi = inputs.spheres_sdf(g, nspheres=8, seed=3)
    phi_t = inputs.spheres_sdf(g, nspheres=8, seed=3, rmin=0.055, rmax=0.155)
    (u, v, w), modes = inputs.fourier_velocity(g, seed=4, return_modes=True)
    dt = inputs.cfl_dt(g, u, v, w)
    hmin, hmax = inputs.spacing_bounds(g)
    blk = api.Block(g)
    log(f"differentiation inputs {n}^3: {time.time() - t0:.1f}s")
    prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4, epsReinitialisation=1e-12)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    ks, es, amps, psis = modes
    heps = 1.5 * hmax
    vol = 1.0 / blk.aj
    with torch.no_grad():
        Ut = autograd.contravariant(blk, *autograd.fourier_velocity(blk, ks, es, dev(amps), psis))
        target = autograd.levelset_trajectory(blk, dev(phi_t), Ut, dt, prm, nsteps, mode="fast",
                                              checkpoint_every=every)
        Ht = autograd.heaviside(target, heps)

    def forward():
        with torch.no_grad():
            U = autograd.contravariant(blk, *autograd.fourier_velocity(blk, ks, es, dev(amps), psis))
            out = autograd.levelset_trajectory(blk, dev(phi), U, dt, prm, nsteps, mode="fast", checkpoint_every=every)
            return (vol * (autograd.heaviside(out, heps) - Ht) ** 2).sum()

    def gradient():
        p = dev(phi).requires_grad_(True)
        a = dev(amps).requires_grad_(True)
        U = autograd.contravariant(blk, *autograd.fourier_velocity(blk, ks, es, a, psis))
        out = autograd.levelset_trajectory(blk, p, U, dt, prm, nsteps, mode="fast", checkpoint_every=every)
        J = (vol * (autograd.heaviside(out, heps) - Ht) ** 2).sum()
        J.backward()
        return p.grad, a.grad
    f_ms = _time(forward, 2, warm=1)
    g_ms = _time(gradient, 2, warm=1)
    return {"workload": f"{n}^3 block, {nsteps} steps x (RK3 + 4 reinitialisation iterations + bcond), checkpoint "
                        f"every {every}, smoothed-Heaviside misfit, gradients w.r.t. phi0 and 16 Fourier amplitudes",
            "forward_ms": round(f_ms, 2), "forward_plus_gradient_ms": round(g_ms, 2),
            "gradient_over_forward": round(g_ms / f_ms, 3),
            "backward_over_forward": round((g_ms - f_ms) / f_ms, 3)}


def load_traffic(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/*.json)."""
    p = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:  # noqa: BLE001
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle-s", type=float, default=2.0,
                    help="untimed steps run for this many seconds before the W warm-up steps, so the "
                         "GPU clocks have come up when the timed region starts (reported in extra)")
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--data-model", default="reference", choices=["folded", "reference"],
                    help="reference: ucn and aj (4 fields per stage), the reference's data model; "
                         "folded: frozen-velocity operator aj*dt*d*ucn (3 fields per stage)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-block batched workload")
    ap.add_argument("--no-diff", action="store_true", help="skip the 192^3 differentiation workload")
    ap.add_argument("--no-inverse", action="store_true", help="skip the 32-channel inverse problem")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--workload", default="all",
                    choices=["all", "step-adjoint", "batch", "differentiation", "consumers", "inverse",
                             "geometric", "fortran-route"],
                    help="run only one secondary workload and print its figures (for profiling); "
                         "'all' is the benchmark line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (scales rocprof "
                         "FETCH_SIZE/WRITE_SIZE to bytes)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nb, "copy_calibration_bytes_written": nb,
                          "launches": 5}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              "(the level-set path would shard by blocks with halo exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import numpy as np
    import torch
    from lsm_amd import api, _lib

    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    L = _lib.load()
    n = args.size
    if args.workload != "all":
        if args.workload == "batch":
            res = bench_batch(args.batch_steps)
        elif args.workload == "differentiation":
            res = bench_differentiation()
        elif args.workload == "inverse":
            res = bench_inverse()
        elif args.workload == "fortran-route":
            res = bench_fortran_route()
        elif args.workload == "consumers":
            g, phi_np, _, _ = make_case(n)
            res = bench_consumers(g, torch.as_tensor(phi_np, device="cuda"), api.Block(g))
        elif args.workload == "geometric":
            g, phi_np, vel, dt = make_case(n)
            res = bench_geometric(g, torch.as_tensor(phi_np, device="cuda"), api.Block(g), vel, dt)
        else:
            g, phi_np, vel, dt = make_case(n)
            blk = api.Block(g)
            ucn = api.rhs_contra_j(blk, *(torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in vel))
            res = bench_step_adjoint(g, torch.as_tensor(phi_np, device="cuda"), ucn, dt, blk)
        print(json.dumps({"workload": args.workload, "result": res}), flush=True)
        return
    t0 = time.time()
    g, phi_np, vel, dt = make_case(n)
    log(f"inputs {n}^3 built in {time.time() - t0:.1f}s (dt = {dt:.3e})")
    blk = api.Block(g)
    phi = torch.as_tensor(phi_np, device="cuda")
    ucn = api.rhs_contra_j(blk, *(torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in vel))
    work = blk.workspace(_lib.LSM_OP_RK3)
    bufs = [phi.clone(), blk.empty()]
    interior = (n - 2) ** 3
    folded = args.data_model == "folded" and args.mode == "fast"
    V = api.fold_velocity(blk, ucn, dt) if folded else None

    def make_step(kind):
        if kind == "folded":
            return lambda src, dst: api.ad_RKTVD3_folded(blk, src, V, out=dst, work=work)
        return lambda src, dst: api.ad_RKTVD3(blk, src, ucn, dt, out=dst, mode=kind, work=work)

    def settle(kind, seconds):
        # clock settle: untimed steps until `seconds` have passed (a short
        # warm-up leaves the first timed steps 10-20 % slow while the clocks
        # come up)
        step = make_step(kind)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < seconds:
            for _ in range(20):
                step(bufs[n % 2], bufs[(n + 1) % 2])
                n += 1
            torch.cuda.synchronize()
        return n, time.perf_counter() - t0

    def timed(kind, steps, warmup, instrument=True):
        step = make_step(kind)
        for s in range(warmup):
            step(bufs[s % 2], bufs[(s + 1) % 2])
        torch.cuda.synchronize()
        base = warmup % 2
        L.lsm_timing_enable(1 if instrument else 0)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        ev0.record()
        for s in range(steps):
            step(bufs[(base + s) % 2], bufs[(base + s + 1) % 2])
        ev1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        kms, kn = C.c_double(), C.c_int64()
        L.lsm_timing_get(_lib.LSM_KT_ADVECT_STAGE, C.byref(kms), C.byref(kn))
        sms, sn = C.c_double(), C.c_int64()
        L.lsm_timing_get(_lib.LSM_KT_ADVECT_SHELL, C.byref(sms), C.byref(sn))
        L.lsm_timing_enable(0)
        return ev0.elapsed_time(ev1), wall, kms, kn, sms, sn, (base + steps) % 2

    head_kind = "folded" if folded else args.mode
    # the K timed steps run without per-kernel events; a second pass of K
    # steps right after, with the library's HIP events around every launch,
    # gives the kernels' average durations
    settle_steps, settle_s = settle(head_kind, args.settle_s) if args.settle_s > 0 else (0, 0.0)
    ms_total, wall, _, _, _, _, last = timed(head_kind, args.steps, args.warmup, instrument=False)
    ms_inst, _, kms, kn, sms, sn, last = timed(head_kind, args.steps, 0, instrument=True)
    base = last

    ms_per_step = ms_total / args.steps
    value = interior * args.steps / (ms_total / 1e3)
    stage_ms = kms.value / max(kn.value, 1)
    # one launch processes every interior cell once; on average 160/3 B each
    alg_bytes_launch = interior * ADV_BYTES_PER_CELL_UPDATE / 3.0
    achieved = alg_bytes_launch / (stage_ms / 1e3) / 1e9
    traffic = load_traffic(f"advect_stage_{n}_{'folded' if folded else args.mode}")
    roofline = {"bound": "hbm", "kernel": "advect_stage", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "avg_launch_ms": round(stage_ms, 5), "launches": int(kn.value),
                "alg_bytes_per_launch": alg_bytes_launch,
                "step_frac_of_roofline": round(value * ADV_BYTES_PER_CELL_UPDATE / 1e9 / HBM_PEAK_GBS, 4)}
    if folded:
        # the folded kernel's own compulsory traffic: stage 1 reads phi^n and
        # V (3) and writes phi (40 B), stages 2-3 also read the stage input (48 B)
        kb = interior * ADV_FOLDED_BYTES_PER_CELL_UPDATE / 3.0
        roofline["kernel_alg_bytes_per_launch"] = kb
        roofline["kernel_frac"] = round(kb / (stage_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step, {value:.3e} cell-updates/s, "
        f"stage kernel {stage_ms:.4f} ms ({achieved:.0f} GB/s alg), shell {sms.value / max(sn.value, 1):.4f} ms")

    extra = {"clock_settle": {"seconds": round(settle_s, 2), "untimed_steps": settle_steps,
                              "note": "untimed steps before the W warm-up steps (GPU clocks up)"},
             "wall_s_timed_region": round(wall, 4), "shell_avg_ms": round(sms.value / max(sn.value, 1), 5),
             "data_model": head_kind, "instrumented_ms_per_step": round(ms_inst / args.steps, 5)}
    if not args.no_extra:
        # the other arithmetic / data-model variants of the same step
        if V is None and args.mode == "fast":
            V = api.fold_velocity(blk, ucn, dt)
        for kind in ("fast", "exact", "folded"):
            if kind == head_kind or (kind == "folded" and V is None):
                continue
            t_ms, _, kms2, kn2, _, _, _ = timed(kind, 10, 2)
            tag = "folded" if kind == "folded" else f"{kind}_refdata"
            extra[f"{tag}_ms_per_step"] = round(t_ms / 10, 5)
            extra[f"{tag}_cell_updates_per_s"] = interior * 10 / (t_ms / 1e3)
            extra[f"{tag}_stage_avg_ms"] = round(kms2.value / max(kn2.value, 1), 5)
    if not args.no_extra:
        # reverse mode of the same step (lsm_rk3_step_adjoint: forward recompute,
        # bcond/WENO3 transposes, dJ/dphi^n and dJ/dU accumulated)
        aw = blk.workspace(_lib.LSM_OP_RK3_ADJOINT)
        lam = bufs[base].clone()
        adj = blk.empty()
        aU = [blk.zeros() for _ in range(3)]
        amode = "fast" if args.mode == "fast" else "exact"
        for _ in range(2):
            api.ad_RKTVD3_adjoint(blk, bufs[base], ucn, dt, lam, mode=amode, adj_ucn=aU, out=adj, work=aw)
        torch.cuda.synchronize()
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        a0.record()
        for _ in range(reps):
            api.ad_RKTVD3_adjoint(blk, bufs[base], ucn, dt, lam, mode=amode, adj_ucn=aU, out=adj, work=aw)
        a1.record()
        torch.cuda.synchronize()
        adj_ms = a0.elapsed_time(a1) / reps
        extra["adjoint_ms_per_step"] = round(adj_ms, 4)
        extra["adjoint_cell_updates_per_s"] = interior / (adj_ms / 1e3)
        extra["adjoint_forward_ratio"] = round(adj_ms / ms_per_step, 3)
        log(f"adjoint: {adj_ms:.3f} ms/step ({adj_ms / ms_per_step:.2f}x the forward step)")
        del aw, lam, adj, aU
    if not args.no_extra:
        # secondary units at the same size: Sussman reinitialisation and water volume
        hmin, hmax = __import__("lsm_amd.inputs", fromlist=["x"]).spacing_bounds(g)
        prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4,
                               epsReinitialisation=1e-12)
        rw = blk.workspace(_lib.LSM_OP_REINIT_ENO2)
        rout = blk.empty()
        api.reinitialisation_test2(blk, bufs[base], prm, out=rout, work=rw)
        torch.cuda.synchronize()
        L.lsm_timing_enable(1)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _, iters = api.reinitialisation_test2(blk, bufs[base], prm, out=rout, work=rw)
        e1.record()
        torch.cuda.synchronize()
        rms, rn = C.c_double(), C.c_int64()
        L.lsm_timing_get(_lib.LSM_KT_REINIT_STAGE, C.byref(rms), C.byref(rn))
        it = int(iters.cpu()[0])
        r_ms = e0.elapsed_time(e1)
        r_stage = rms.value / max(rn.value, 1)
        extra["reinit_cell_iters_per_s"] = (n ** 3) * it / (r_ms / 1e3)
        extra["reinit_iterations"] = it
        extra["reinit_stage_avg_ms"] = round(r_stage, 5)
        extra["reinit_stage_alg_GBps"] = round((n ** 3) * REINIT_BYTES_PER_CELL_ITER / 3.0 / (r_stage / 1e3) / 1e9, 1)
        vw = blk.workspace(_lib.LSM_OP_WATER_VOLUME)
        api.TotalWaterVolume(blk, bufs[base], with_mask=False, work=vw)
        torch.cuda.synchronize()
        e0.record()
        vol, _ = api.TotalWaterVolume(blk, bufs[base], with_mask=False, work=vw)
        e1.record()
        torch.cuda.synchronize()
        v_ms = e0.elapsed_time(e1)
        extra["volume_cells_per_s"] = ((n - 1) ** 3) / (v_ms / 1e3)
        extra["volume_ms"] = round(v_ms, 4)
        extra["volume_alg_GBps"] = round(((n - 1) ** 3) * VOLUME_BYTES_PER_CELL / (v_ms / 1e3) / 1e9, 1)
        L.lsm_timing_enable(0)
        log(f"reinit: {it} iterations in {r_ms:.2f} ms (stage {r_stage:.4f} ms), volume {v_ms:.3f} ms")
        # one whole levelsetmethod call (volume, velocity, RK3, 4 Sussman
        # iterations, bcond) from the velocity field
        uvw = [torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in vel]
        pl = bufs[base].clone()
        pn = pl.clone()
        lw = blk.workspace(_lib.LSM_OP_LEVELSET_STEP)
        api.levelsetmethod(blk, pl, pn, *uvw, prm, dt, ntime=1, work=lw)
        torch.cuda.synchronize()
        e0.record()
        reps = 3
        for r in range(reps):
            api.levelsetmethod(blk, pl, pn, *uvw, prm, dt, ntime=1 + r, work=lw)
        e1.record()
        torch.cuda.synchronize()
        ls_ms = e0.elapsed_time(e1) / reps
        extra["levelsetmethod_ms"] = round(ls_ms, 3)
        extra["levelsetmethod_cells_per_s"] = interior / (ls_ms / 1e3)
        log(f"levelsetmethod (4 reinit iterations): {ls_ms:.2f} ms per call")

    if not args.no_extra:
        extra["geometric"] = bench_geometric(g, torch.as_tensor(phi_np, device="cuda"), blk, vel, dt)
        log(f"geometric: {extra['geometric']}")
    if not args.no_extra:
        extra["consumers"] = bench_consumers(g, bufs[base], blk)
        log(f"consumers: {extra['consumers']}")
    if not args.no_extra:
        # on the initial 8-sphere level set (the step reinitialises every
        # time, so its input stays close to a distance function; the advected
        # field of the timed steps was never reinitialised)
        extra.update(bench_step_adjoint(g, torch.as_tensor(phi_np, device="cuda"), ucn, dt, blk))
        log(f"full step: fwd {extra['step_fwd_ms']:.3f} ms, adjoint {extra['step_adjoint_ms']:.3f} ms "
            f"({extra['step_adjoint_forward_ratio']:.2f}x)")
    del work, V
    torch.cuda.empty_cache()
    if not args.no_extra and not args.no_batch:
        extra["batch"] = bench_batch(args.batch_steps)
        torch.cuda.empty_cache()
    if not args.no_extra and not args.no_diff:
        extra["differentiation"] = bench_differentiation()
        torch.cuda.empty_cache()
    if not args.no_extra:
        try:
            extra["fortran_route"] = bench_fortran_route()
            log(f"fortran route: {extra['fortran_route']}")
        except Exception as e:  # noqa: BLE001
            log(f"fortran route not timed: {e}")
        torch.cuda.empty_cache()
    if not args.no_extra and not args.no_inverse:
        extra["inverse"] = bench_inverse()
        log(f"inverse: forward {extra['inverse']['batched_forward_ms']} ms, forward + gradient "
            f"{extra['inverse']['forward_plus_gradient_ms']} ms")
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu_baseline:
        t1 = time.time()
        cpu = cpu_baseline()
        log(f"cpu baseline ({cpu['kind']}): {cpu['value']:.3e} cell-updates/s in {time.time() - t1:.1f}s")
        if "geometric" in extra:
            t1 = time.time()
            geo_cpu = cpu_geometric_baseline()
            if geo_cpu:
                extra["geometric"]["cpu_reference"] = geo_cpu
                log(f"geometric cpu baseline: {geo_cpu['value']:.3e} band-nodes/s in {time.time() - t1:.1f}s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"forward ad_RKTVD3 (WENO3 + SSP-RK3) on one {n}^3 curvilinear block, "
                               "8-sphere SDF, 16-mode divergence-free Fourier velocity, CFL 0.4",
                   "grid": [n, n, n], "interior_cells": interior, "arithmetic": args.mode},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
