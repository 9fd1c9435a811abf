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
REINIT_BYTES_PER_CELL_ITER = 304.0


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
                        "sample": sample + f", compiled reference Fortran (amdflang -O2), {ranks} MPI "
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--data-model", default="reference", choices=["folded", "reference"],
                    help="reference: ucn and aj (4 fields per stage), the reference's data model; "
                         "folded: frozen-velocity operator aj*dt*d*ucn (3 fields per stage)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
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

    def timed(kind, steps, warmup):
        step = make_step(kind)
        for s in range(warmup):
            step(bufs[s % 2], bufs[(s + 1) % 2])
        torch.cuda.synchronize()
        base = warmup % 2
        L.lsm_timing_enable(1)
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
    ms_total, wall, kms, kn, sms, sn, last = timed(head_kind, args.steps, args.warmup)
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

    extra = {"wall_s_timed_region": round(wall, 4), "shell_avg_ms": round(sms.value / max(sn.value, 1), 5),
             "data_model": head_kind}
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

    cpu = None
    if not args.no_cpu_baseline:
        t1 = time.time()
        cpu = cpu_baseline()
        log(f"cpu baseline ({cpu['kind']}): {cpu['value']:.3e} cell-updates/s in {time.time() - t1:.1f}s")

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
