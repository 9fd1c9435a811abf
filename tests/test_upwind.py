"""orderAD = 0: the first-order upwind right-hand side (calc_RH_AD) and the
RK-TVD3 step built on it (ad_RKTVD3.F90:57).

CPU: the oracle against golden vectors of the compiled reference (bit for
bit) and against the closed form (a linear field under a constant velocity is
differentiated exactly).  GPU: the HIP entry points against the oracle, bit
for bit, with masks and every boundary type."""
import numpy as np
import pytest

import oracle as O
from fixtures import load
from lsm_amd import inputs


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    diff = int(np.sum(a != b))
    assert diff == 0, f"{diff} of {a.size} values differ, max |d| = {np.max(np.abs(a - b))}"


@pytest.fixture(scope="module")
def up():
    return load("upwind")


def _blk(fx, btype=None):
    return O.Block(fx.grid, btype=btype or fx.par("btype"), big_phi=fx.par("big_phi"))


@pytest.mark.parametrize("key,mask", [("rhs", False), ("rhs_mask", True)])
def test_oracle_upwind_rhs_golden(up, key, mask):
    blk = _blk(up)
    U = O.contra(blk, *up.velocity)
    rh = O.upwind_rhs(blk, up.inp("phi"), U, up.inp("mask") if mask else None)
    _eq(rh, up.out(key))


@pytest.mark.parametrize("key,mask", [("advect", False), ("advect_mask", True), ("advect_iter2", False)])
def test_oracle_upwind_rk3_golden(up, key, mask):
    bt = up.kw(key, "btype") or up.par("btype")
    blk = _blk(up, bt)
    U = O.contra(blk, *up.velocity)
    numiter = up.kw(key, "numiter_lsm", 1)
    m = up.inp("mask") if mask else None
    dt = up.par("delti_lsm")
    _eq(O.rk3(blk, up.inp("phi"), U, dt, m, nsteps=1, numiter=numiter, order_ad=0), up.out(key, 0))
    _eq(O.rk3(blk, up.inp("phi"), U, dt, m, nsteps=up.kw(key, "nsteps"), numiter=numiter, order_ad=0),
        up.out(key, 1))


def test_oracle_upwind_rhs_of_linear_field_is_exact():
    """On a uniform grid, U^d (dphi/dxi_d) summed over d for phi linear: every
    one-sided and upwind difference is exact."""
    g = inputs.uniform_grid(9, 8, 7, h=0.5)
    a = (0.3, -1.1, 0.7)
    phi = a[0] * g.x + a[1] * g.y + a[2] * g.z
    rng = np.random.default_rng(2)
    U = [rng.standard_normal(g.shape) for _ in range(3)]
    rh = O.upwind_rhs(O.Block(g), phi, U)
    # dc = de = dz = 1 (index space): d phi / d xi_d = a_d * h
    ref = sum(U[d] * (a[d] * 0.5) for d in range(3))
    np.testing.assert_allclose(rh, ref, rtol=0, atol=1e-13)


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("mask", [False, True])
def test_gpu_upwind_rhs_golden(up, mask):
    import torch
    from lsm_amd import api
    print(f"[upwind] rhs golden mask={mask}", flush=True)
    blk = api.Block(up.grid, btype=up.par("btype"), big_phi=up.par("big_phi"))
    U = api.rhs_contra_j(blk, *[_dev(a) for a in up.velocity])
    m = _dev(up.inp("mask"), torch.int32) if mask else None
    rh = api.calc_RH_AD(blk, _dev(up.inp("phi")), U, m, out=blk.zeros())
    _eq(_host(rh), up.out("rhs_mask" if mask else "rhs"))


@pytest.mark.gpu
@pytest.mark.parametrize("key,mask", [("advect", False), ("advect_mask", True), ("advect_iter2", False)])
def test_gpu_upwind_rk3_golden(up, key, mask):
    import torch
    from lsm_amd import api
    print(f"[upwind] rk3 golden {key}", flush=True)
    bt = up.kw(key, "btype") or up.par("btype")
    blk = api.Block(up.grid, btype=bt, big_phi=up.par("big_phi"))
    U = api.rhs_contra_j(blk, *[_dev(a) for a in up.velocity])
    m = _dev(up.inp("mask"), torch.int32) if mask else None
    numiter = up.kw(key, "numiter_lsm", 1)
    cur = _dev(up.inp("phi"))
    for s in range(up.kw(key, "nsteps")):
        for _ in range(numiter):
            cur = api.ad_RKTVD3(blk, cur, U, up.par("delti_lsm"), mode="upwind", adv_mask=m)
        if s == 0:
            _eq(_host(cur), up.out(key, 0))
    _eq(_host(cur), up.out(key, 1))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(67, 45, 38), (64, 64, 64)])
def test_gpu_upwind_matches_oracle(shape):
    """Larger seeded blocks: 3 upwind RK3 steps equal the oracle bit for bit."""
    from lsm_amd import api
    ni, nj, nk = shape
    print(f"[upwind] oracle parity {shape}", flush=True)
    g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=8, nj=nj, nk=nk)
    phi = inputs.spheres_sdf(g, nspheres=5, seed=9, rmin=0.1, rmax=0.3)
    u, v, w = inputs.fourier_velocity(g, seed=10)
    dt = inputs.cfl_dt(g, u, v, w)
    oblk = O.Block(g, btype=(1, 0, 2, 1, 1, 2), big_phi=0.2)
    blk = api.Block(g, btype=(1, 0, 2, 1, 1, 2), big_phi=0.2)
    U = api.rhs_contra_j(blk, *[_dev(a) for a in (u, v, w)])
    ref = O.rk3(oblk, phi, [_host(t) for t in U], dt, nsteps=3, order_ad=0)
    cur = _dev(phi)
    for _ in range(3):
        cur = api.ad_RKTVD3(blk, cur, U, dt, mode="upwind")
    _eq(_host(cur), ref)
