"""Multi-instance execution (lsm_multi_begin .. lsm_multi_end): the recorded
launch sequences of several blocks of different shapes merged into one launch
per kernel give, bit for bit, what the same calls give one after another --
the forward step, the recorded step and its reverse sweep.  Misuse (different
sequences, entry points that cannot be recorded, missing workspaces) is
rejected with an error and leaves the library usable."""
import ctypes as C

import numpy as np
import pytest

SHAPES = [(24, 18, 20), (40, 12, 33), (17, 29, 26)]


def test_multi_state_errors_without_gpu():
    from lsm_amd import _lib
    L = _lib.load()
    assert L.lsm_multi_instance() == _lib.LSM_ERR_ARG
    assert L.lsm_multi_end(None) == _lib.LSM_ERR_ARG
    assert L.lsm_multi_begin() == _lib.LSM_OK
    assert L.lsm_multi_begin() == _lib.LSM_ERR_ARG
    assert L.lsm_multi_cancel() == _lib.LSM_OK
    assert L.lsm_multi_begin() == _lib.LSM_OK
    assert L.lsm_multi_end(None) == _lib.LSM_OK     # nothing recorded: nothing to launch


def _cases():
    import torch
    from lsm_amd import api, inputs
    out = []
    for q, (ni, nj, nk) in enumerate(SHAPES):
        g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=40 + q, nj=nj, nk=nk)
        phi = inputs.spheres_sdf(g, nspheres=3, seed=50 + q, rmin=0.15, rmax=0.3) * (0.7 + (g.x - 0.4) ** 2)
        u, v, w = inputs.fourier_velocity(g, seed=60 + q)
        blk = api.Block(g)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
        ucn = api.rhs_contra_j(blk, dev(u), dev(v), dev(w))
        out.append((blk, dev(phi), ucn, inputs.cfl_dt(g, u, v, w)))
    return out


def _params():
    from lsm_amd import api
    return api.ReinitParams(epslsm=0.08, RNtimestep=0.01, RNitermax=3, epsReinitialisation=1e-12)


def _eq(a, b):
    import torch
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.array_equal(a, b), f"{int(np.sum(a != b))} values differ"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_multi_forward_step_matches_single(mode):
    """ad_RKTVD3 + reinitialisation_test2 + bcond_lsm recorded for three
    blocks and run as merged launches equal the calls made one by one."""
    import torch
    from lsm_amd import api
    print(f"[multi] forward step {mode}", flush=True)
    cases = _cases()
    prm = _params()
    ref = []
    for blk, phi, ucn, dt in cases:
        x = api.ad_RKTVD3(blk, phi, ucn, dt, mode=mode)
        y, _ = api.reinitialisation_test2(blk, x, prm)
        ref.append(api.bcond_lsm(blk, y).clone())
    outs = []
    with api.multi() as m:
        for blk, phi, ucn, dt in cases:
            m.instance()
            x = api.ad_RKTVD3(blk, phi, ucn, dt, mode=mode)
            y, _ = api.reinitialisation_test2(blk, x, prm)
            outs.append(api.bcond_lsm(blk, y))
    assert m.count == len(cases)
    for a, b in zip(outs, ref):
        _eq(a, b)


@pytest.mark.gpu
def test_multi_recorded_step_and_reverse_sweep_match_single():
    """lsm_levelset_step_record and lsm_levelset_step_adjoint_recorded over
    three blocks, merged: phi^(n+1), dJ/dphi^n and dJ/d(U/J) bit for bit."""
    import torch
    from lsm_amd import api
    print("[multi] recorded step + reverse sweep", flush=True)
    cases = _cases()
    prm = _params()
    rng = np.random.default_rng(7)
    lams = [torch.as_tensor(rng.standard_normal(blk.shape), device="cuda") for blk, _, _, _ in cases]
    ref = []
    for (blk, phi, ucn, dt), lam in zip(cases, lams):
        p1, tr = api.levelset_step_record(blk, phi, ucn, dt, prm, mode="fast")
        aU = [blk.zeros() for _ in range(3)]
        adj = api.levelset_step_adjoint_recorded(blk, phi, ucn, dt, prm, tr, lam, mode="fast", adj_ucn=aU)
        ref.append((p1.clone(), adj.clone(), [t.clone() for t in aU]))
    recs = []
    with api.multi() as m:
        for blk, phi, ucn, dt in cases:
            m.instance()
            recs.append(api.levelset_step_record(blk, phi, ucn, dt, prm, mode="fast"))
    outs = []
    with api.multi() as m:
        for (blk, phi, ucn, dt), lam, (p1, tr) in zip(cases, lams, recs):
            m.instance()
            aU = [blk.zeros() for _ in range(3)]
            outs.append((p1, api.levelset_step_adjoint_recorded(blk, phi, ucn, dt, prm, tr, lam, mode="fast",
                                                                adj_ucn=aU), aU))
    for (p1, adj, aU), (r1, radj, raU) in zip(outs, ref):
        _eq(p1, r1)
        _eq(adj, radj)
        for a, b in zip(aU, raU):
            _eq(a, b)


@pytest.mark.gpu
def test_multi_rejects_misuse_and_recovers():
    import torch
    from lsm_amd import api, _lib
    from lsm_amd._lib import LsmError
    print("[multi] misuse", flush=True)
    cases = _cases()
    (b0, p0, u0, dt0), (b1, p1, u1, dt1) = cases[0], cases[1]
    # different kernel sequences (fast / exact arithmetic)
    with pytest.raises(LsmError, match="differs"):
        with api.multi() as m:
            m.instance()
            api.ad_RKTVD3(b0, p0, u0, dt0, mode="fast")
            m.instance()
            api.ad_RKTVD3(b1, p1, u1, dt1, mode="exact")
    # an entry point that cannot be recorded
    with pytest.raises(LsmError, match="cannot be recorded"):
        with api.multi() as m:
            m.instance()
            api.TotalWaterVolume(b0, p0)
    # no workspace: the library's own buffer would be shared by the instances
    L = _lib.load()
    out = b0.empty()
    assert L.lsm_multi_begin() == 0 and L.lsm_multi_instance() == 0
    rc = L.lsm_rk3_step(b0.ref, C.c_void_p(out.data_ptr()), C.c_void_p(p0.data_ptr()), None,
                        *[C.c_void_p(t.data_ptr()) for t in u0], dt0, 0, None, 0, None)
    assert rc == _lib.LSM_ERR_ARG and b"own workspace" in L.lsm_last_error()
    assert L.lsm_multi_cancel() == 0
    # the library is usable afterwards and launches immediately again
    x = api.ad_RKTVD3(b0, p0, u0, dt0, mode="fast")
    with api.multi() as m:
        m.instance()
        y = api.ad_RKTVD3(b0, p0, u0, dt0, mode="fast")
    _eq(x, y)


@pytest.mark.gpu
def test_multi_trajectory_matches_per_block_trajectories():
    """Five checkpointed steps of three blocks advanced together
    (multi_trajectory) against levelset_trajectory per block: the final
    fields and the gradients with respect to phi0 and U/J bit for bit."""
    import torch
    from lsm_amd import autograd
    print("[multi] trajectory", flush=True)
    cases = _cases()
    prm = _params()
    dt = min(c[3] for c in cases)
    rng = np.random.default_rng(11)
    ws = [torch.as_tensor(rng.standard_normal(blk.shape), device="cuda") for blk, _, _, _ in cases]

    def inputs():
        phis = [p.clone().requires_grad_(True) for _, p, _, _ in cases]
        us = [[u.clone().requires_grad_(True) for u in ucn] for _, _, ucn, _ in cases]
        return phis, us

    phis, us = inputs()
    ref = []
    for (blk, _, _, _), p, u, w in zip(cases, phis, us, ws):
        out = autograd.levelset_trajectory(blk, p, u, dt, prm, 5, mode="fast", checkpoint_every=2)
        (out * w).sum().backward()
        ref.append((out.detach().clone(), p.grad.clone(), [t.grad.clone() for t in u]))
    phis, us = inputs()
    outs = autograd.multi_trajectory([c[0] for c in cases], phis, us, dt, prm, 5, mode="fast",
                                     checkpoint_every=2)
    sum((o * w).sum() for o, w in zip(outs, ws)).backward()
    for b in range(len(cases)):
        _eq(outs[b].detach(), ref[b][0])
        _eq(phis[b].grad, ref[b][1])
        for c in range(3):
            _eq(us[b][c].grad, ref[b][2][c])
