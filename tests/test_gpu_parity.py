"""GPU parity: the HIP library (through its C ABI) against the golden vectors
of the compiled reference Fortran and against the CPU oracle.

Tolerances (written per test):
  * LSM_ADVECT_EXACT arithmetic and the reinitialisation/volume kernels keep
    the reference operation order, so most results are bit-identical; where
    device libm (sin/cos) or reductions can differ by an ulp the bound is
    1e-13 relative to max|phi|.
  * LSM_ADVECT_FAST (one division per WENO3 face value): 1e-12 max|phi| per
    step, 1e-9 after 100 steps (the target of the port).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
from fixtures import load  # noqa: E402
from lsm_amd import api, inputs  # noqa: E402


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def linf_rel(a, b):
    b = np.asarray(b)
    return float(np.max(np.abs(np.asarray(a) - b))) / max(float(np.max(np.abs(b))), 1e-300)


@pytest.fixture(scope="module")
def curv():
    fx = load("curvilinear")
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    u, v, w = (dev(a) for a in fx.velocity)
    ucn = api.rhs_contra_j(blk, u, v, w)
    return fx, blk, ucn


def test_contra_velocity_bitwise(curv):
    fx, blk, ucn = curv
    for c in range(3):
        assert np.array_equal(host(ucn[c]), fx.out("contra", c))


@pytest.mark.parametrize("mode,tol", [("exact", 0.0), ("fast", 1e-13)])
def test_weno3_rhs(curv, mode, tol):
    fx, blk, ucn = curv
    rh = host(api.calc_RH_AD_WENO3(blk, dev(fx.inp("phi")), ucn, mode=mode))
    ref = fx.out("weno_rhs")
    assert linf_rel(rh, ref) <= tol, linf_rel(rh, ref)


def test_bcond_bitwise(curv):
    fx, blk, _ = curv
    p = dev(fx.inp("phi"))
    api.bcond_lsm(blk, p)
    assert np.array_equal(host(p), fx.out("bcond"))


@pytest.mark.parametrize("mode,tol1,tol3", [("exact", 0.0, 0.0), ("fast", 1e-12, 3e-12)])
def test_rk3_steps(curv, mode, tol1, tol3):
    fx, blk, ucn = curv
    dt = fx.par("delti_lsm")
    p = dev(fx.inp("phi"))
    outs = []
    for _ in range(3):
        p = api.ad_RKTVD3(blk, p, ucn, dt, mode=mode)
        outs.append(host(p))
    assert linf_rel(outs[0], fx.out("advect", 0)) <= tol1
    assert linf_rel(outs[2], fx.out("advect", 1)) <= tol3


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("eno", [0, 1])
def test_normgrad(curv, order, eno):
    fx, blk, _ = curv
    ng = host(api.calc_NormGrad_d(blk, dev(fx.inp("phi")), dev(fx.inp("phi0")), order, eno))
    assert np.array_equal(ng, fx.out(f"normgrad_o{order}_e{eno}"))


@pytest.mark.parametrize("order", [1, 2])
def test_reinit_eno2(curv, order):
    fx, blk, _ = curv
    prm = api.ReinitParams(epslsm=fx.par("epslsm"), RNtimestep=fx.par("rn_timestep"), RNitermax=5,
                           epsReinitialisation=fx.par("eps_reinit"),
                           OrderReinitialisationBoundaries=order)
    out, it = api.reinitialisation_test2(blk, dev(fx.inp("phi0")), prm)
    ref = fx.out(f"reinit_eno2_o{order}")
    # device sin/cos may differ from glibc by an ulp inside the interface band
    assert linf_rel(host(out), ref) <= 1e-13
    _, it_ref = O.reinit_eno2(O.Block(fx.grid), fx.inp("phi0"), fx.par("epslsm"),
                              fx.par("eps_reinit"), fx.par("rn_timestep"), 5, order)
    assert int(host(it)[0]) == it_ref


@pytest.mark.parametrize("shape,order,eno", [((65, 9, 8), 2, True), ((66, 7, 9), 1, False),
                                              ((67, 6, 7), 2, True), ((130, 5, 6), 3, False),
                                              ((4, 9, 7), 2, False)])
def test_reinit_ragged_extents(shape, order, eno):
    """Extents where the i = ni face node falls in lane 0, 1 or 2 of a tile (the
    first one keeps the i faces in the boundary shell; the other two take the
    one-sided i stencil from the tile's halo columns), a two-tile row and a
    block too narrow for the tiles' i faces match the oracle's
    reinitialisation for every boundary order (1e-13: device sin/cos)."""
    ni, nj, nk = shape
    g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=91, nj=nj, nk=nk)
    phi = inputs.spheres_sdf(g, nspheres=3, seed=92, rmin=0.2, rmax=0.35) * (0.5 + (g.x - 0.3) ** 2)
    lo, hi = inputs.spacing_bounds(g)
    prm = api.ReinitParams(epslsm=1.5 * hi, RNtimestep=0.5 * lo, RNitermax=4, epsReinitialisation=1e-12,
                           OrderReinitialisationBoundaries=order, ENOBCReinitialisation=eno)
    out, it = api.reinitialisation_test2(api.Block(g), dev(phi), prm)
    ref, it_ref = O.reinit_eno2(O.Block(g), phi, 1.5 * hi, 1e-12, 0.5 * lo, 4, order, eno)
    assert int(host(it)[0]) == it_ref
    assert linf_rel(host(out), ref) <= 1e-13, (shape, linf_rel(host(out), ref))


@pytest.mark.parametrize("m", [0, 1])
def test_sign_ini(curv, m):
    fx, blk, _ = curv
    s, g = api.calc_sign_ini(blk, dev(fx.inp("phi0")), fx.par("epslsm"), m)
    assert linf_rel(host(s), fx.out(f"sign_ini_m{m}", 0)) <= 1e-15
    assert np.array_equal(host(g), fx.out(f"sign_ini_m{m}", 1))


def test_rh_rn_and_lambda(curv):
    fx, blk, _ = curv
    rh = host(api.calc_RH_RN(blk, dev(fx.inp("phi")), dev(fx.inp("sign"))))
    assert np.array_equal(rh, fx.out("rh_rn"))
    lam = host(api.calc_lambda_RN(blk, dev(fx.inp("phi0")), dev(fx.inp("ng0")), dev(fx.inp("phi")),
                                  fx.par("rn_timestep"), fx.par("epslsm")))
    assert linf_rel(lam, fx.out("lambda_rn")) <= 1e-13


@pytest.mark.parametrize("key,method", [("lambda_rn", 1), ("lambda_rn_m0", 0)])
def test_lambda_filtered(key, method):
    """calc_lambda_RN with sussman_correction_method = 1 (testfilter) against
    the compiled reference; 1e-13 as the pointwise lambda (device cos in
    deltaf)."""
    fx = load("simpson")
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    lam = host(api.calc_lambda_RN(blk, dev(fx.inp("phi0")), dev(fx.inp("ng0")), dev(fx.inp("phi")),
                                  fx.kw(key, "lambda_dt"), fx.par("epslsm"), sussman_correction_method=method))
    assert linf_rel(lam, fx.out(key)) <= 1e-13


@pytest.mark.parametrize("m", [0, 1])
@pytest.mark.parametrize("t", [0, 1])
def test_reinit_godunov_filtered_lambda(m, t):
    fx = load("simpson")
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    prm = api.ReinitParams(epslsm=fx.par("epslsm"), RNtimestep=fx.par("rn_timestep"), RNitermax=4,
                           M_Sphi0=m, M_TI=t, sussman_correction_method=1)
    out = host(api.reinitialization_benchmark(blk, dev(fx.inp("phi0")), prm))
    assert linf_rel(out, fx.out(f"reinit_godunov_m{m}_t{t}")) <= 1e-13


@pytest.mark.parametrize("m", [0, 1])
@pytest.mark.parametrize("t", [0, 1])
def test_reinit_godunov(curv, m, t):
    fx, blk, _ = curv
    prm = api.ReinitParams(epslsm=fx.par("epslsm"), RNtimestep=fx.par("rn_timestep"), RNitermax=4,
                           M_Sphi0=m, M_TI=t)
    out = host(api.reinitialization_benchmark(blk, dev(fx.inp("phi0")), prm))
    assert linf_rel(out, fx.out(f"reinit_godunov_m{m}_t{t}")) <= 1e-13


def test_water_volume(curv):
    fx, blk, _ = curv
    vol, mask = api.TotalWaterVolume(blk, dev(fx.inp("phi")))
    ref = fx.out("volume", 0)
    assert np.max(np.abs(host(vol) - ref)) <= 1e-13 * np.sum(ref)
    assert np.array_equal(host(mask), fx.out("volume", 1))


@pytest.mark.parametrize("plane", ["flat", "tilt", "diag", "sphere"])
def test_water_volume_cases(plane):
    fx = load("volume")
    blk = api.Block(fx.grid)
    p = dev(fx.inp(f"plane_{plane}"))
    vol, mask = api.TotalWaterVolume(blk, p)
    ref = fx.out(f"volume_{plane}", 0)
    assert np.max(np.abs(host(vol) - ref)) <= 1e-13 * max(np.sum(ref), 1e-30)
    assert np.array_equal(host(mask), fx.out(f"volume_{plane}", 1))
    assert np.array_equal(host(api.calc_NormGrad_d(blk, p, p)), fx.out(f"normgrad_{plane}"))


@pytest.mark.parametrize("name", ["curvilinear", "channel"])
@pytest.mark.parametrize("mode,tol", [("exact", 1e-13), ("fast", 1e-12)])
def test_levelsetmethod_steps(name, mode, tol):
    fx = load(name)
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    u, v, w = (dev(a) for a in fx.velocity)
    prm = api.ReinitParams(epslsm=fx.par("epslsm"), RNtimestep=fx.par("rn_timestep"),
                           RNitermax=fx.kw("levelset", "rn_itermax"),
                           epsReinitialisation=fx.par("eps_reinit"))
    phi = dev(fx.inp("phi"))
    vols = []
    for s in range(fx.kw("levelset", "nsteps")):
        phi_n = phi.clone()
        vol = api.levelsetmethod(blk, phi, phi_n, u, v, w, prm, fx.par("delti_lsm"), ntime=1 + s,
                                 mode=mode)
        vols.append(float(host(vol).sum()))
    assert linf_rel(host(phi), fx.out("levelset")) <= tol
    gv = fx.d["out_levelset_volume"]
    assert np.max(np.abs(np.array(vols) - gv)) <= 1e-12 * np.max(gv)


def test_boundary_types_and_mask():
    fx = load("boundary")
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    phi = fx.inp("phi")
    p = dev(phi)
    api.bcond_lsm(blk, p)
    assert np.array_equal(host(p), fx.out("bcond"))
    for key, bt, big in (("bcond_b1", (1,) * 6, fx.par("big_phi")), ("bcond_b2", (2,) * 6, 0.2)):
        b2 = api.Block(fx.grid, btype=bt, big_phi=big)
        p = dev(phi)
        api.bcond_lsm(b2, p)
        assert np.array_equal(host(p), fx.out(key)), key
    u, v, w = (dev(a) for a in fx.velocity)
    ucn = api.rhs_contra_j(blk, u, v, w)
    mask = dev(fx.inp("mask"), torch.int32)
    rh = host(api.calc_RH_AD_WENO3(blk, dev(phi), ucn, adv_mask=mask, mode="exact"))
    assert np.array_equal(rh, fx.out("weno_rhs_mask"))
    dt = fx.par("delti_lsm")
    for key, m in (("advect_mask", mask), ("advect", None)):
        q = dev(phi)
        for _ in range(2):
            q = api.ad_RKTVD3(blk, q, ucn, dt, adv_mask=m, mode="exact")
        assert np.array_equal(host(q), fx.out(key, 1)), key


# ---------------------------------------------------------------------------
# larger grids against the CPU oracle
# ---------------------------------------------------------------------------

def _case(n, seed, amp=0.03):
    g = inputs.sinusoidal_cube_grid(n, amp=amp, seed=seed, nj=n + 3, nk=n - 2)
    phi = inputs.spheres_sdf(g, nspheres=8, seed=seed + 1)
    u, v, w = inputs.fourier_velocity(g, seed=seed + 2)
    return g, phi, (u, v, w), inputs.cfl_dt(g, u, v, w)


@pytest.mark.parametrize("mode,tol", [("exact", 0.0), ("fast", 1e-12)])
def test_rk3_vs_oracle_48(mode, tol):
    g, phi, vel, dt = _case(48, 21)
    blk = api.Block(g)
    ob = O.Block(g)
    ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
    U = O.contra(ob, *vel)
    p = dev(phi)
    ref = phi
    for step in range(4):
        p = api.ad_RKTVD3(blk, p, ucn, dt, mode=mode)
        ref = O.rk3(ob, ref, U, dt)
        err = linf_rel(host(p), ref)
        assert err <= tol * (step + 1), (step, err)


def test_rk3_100_steps_fast_vs_oracle():
    """100 steps of the fast arithmetic stay within 1e-9 max|phi| of the reference order."""
    g, phi, vel, dt = _case(24, 31)
    blk = api.Block(g)
    ob = O.Block(g)
    ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
    U = O.contra(ob, *vel)
    p = dev(phi)
    for _ in range(100):
        p = api.ad_RKTVD3(blk, p, ucn, dt, mode="fast")
    ref = O.rk3(ob, phi, U, dt, nsteps=100)
    assert linf_rel(host(p), ref) <= 1e-9
    # identical wet/dry sign pattern
    assert np.array_equal(host(p) > 0, ref > 0)


def test_reinit_and_volume_vs_oracle_40():
    g, phi, vel, dt = _case(40, 41)
    hmin, hmax = inputs.spacing_bounds(g)
    phi0 = phi * (0.1 + (g.x - 0.5) ** 2)
    blk = api.Block(g)
    ob = O.Block(g)
    prm = api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=6)
    out, it = api.reinitialisation_test2(blk, dev(phi0), prm)
    ref, it_ref = O.reinit_eno2(ob, phi0, 1.5 * hmax, 1e-5, 0.5 * hmin, 6)
    assert linf_rel(host(out), ref) <= 1e-13
    assert int(host(it)[0]) == it_ref
    gd = host(api.reinitialization_benchmark(blk, dev(phi0), api.ReinitParams(
        epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=6, M_Sphi0=0)))
    assert linf_rel(gd, O.reinit_godunov(ob, phi0, 1.5 * hmax, 0.5 * hmin, 6, 0, 1)) <= 1e-13
    vol, mask = api.TotalWaterVolume(blk, dev(phi))
    vref, mref = O.water_volume(ob, phi)
    assert abs(float(host(vol).sum()) - vref.sum()) <= 1e-10 * vref.sum()
    assert np.array_equal(host(mask), mref)


# ---------------------------------------------------------------------------
# full size (256^3): size-independent properties
# ---------------------------------------------------------------------------

def test_linear_field_exact_transport_256():
    """WENO3 reproduces a linear field exactly, so with a constant velocity one
    RK3 step moves phi = a.x + b by -dt u.a at every node at least seven nodes
    from the boundary (boundary extrapolation reaches two nodes per stage)."""
    n = 256
    g = inputs.uniform_grid(n, n, n, h=1.0 / (n - 1))
    a = np.array([0.3, -0.7, 0.55])
    phi = a[0] * g.x + a[1] * g.y + a[2] * g.z + 0.1
    uvw = (np.full(g.shape, 0.8), np.full(g.shape, -0.35), np.full(g.shape, 0.5))
    dt = inputs.cfl_dt(g, *uvw)
    blk = api.Block(g)
    ucn = api.rhs_contra_j(blk, *(dev(x) for x in uvw))
    out = host(api.ad_RKTVD3(blk, dev(phi), ucn, dt, mode="fast"))
    expect = phi - dt * (0.8 * a[0] - 0.35 * a[1] + 0.5 * a[2])
    s = slice(7, -7)
    err = np.max(np.abs(out[s, s, s] - expect[s, s, s]))
    assert err <= 1e-12, err
    del out, phi, expect


def test_fast_vs_exact_256_and_volume_conservation():
    """At 256^3 the fast arithmetic stays within 1e-12 max|phi| of the
    reference-order arithmetic after one step, and the water volume of the
    spheres moves by less than the second-order truncation bound."""
    g, phi, vel, dt = None, None, None, None
    g = inputs.sinusoidal_cube_grid(256, seed=2)
    phi = inputs.spheres_sdf(g, seed=3)
    vel = inputs.fourier_velocity(g, seed=4)
    dt = inputs.cfl_dt(g, *vel)
    blk = api.Block(g)
    ucn = api.rhs_contra_j(blk, *(dev(x) for x in vel))
    p0 = dev(phi)
    pf = api.ad_RKTVD3(blk, p0, ucn, dt, mode="fast")
    pe = api.ad_RKTVD3(blk, p0, ucn, dt, mode="exact")
    d = float((pf - pe).abs().max() / pe.abs().max())
    assert d <= 1e-12, d
    v0, _ = api.TotalWaterVolume(blk, p0, with_mask=False)
    v1, _ = api.TotalWaterVolume(blk, pf, with_mask=False)
    v0 = float(v0.sum())
    v1 = float(v1.sum())
    assert abs(v1 - v0) <= 1e-3 * v0


@pytest.mark.parametrize("shape", [(65, 33, 20), (66, 9, 11), (17, 13, 5), (129, 7, 4)])
def test_rk3_ragged_extents_exact(shape):
    """Extents that leave partial tiles in i and j and short k chunks give the
    reference result bit for bit (exact arithmetic)."""
    ni, nj, nk = shape
    g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=51, nj=nj, nk=nk)
    phi = inputs.spheres_sdf(g, nspheres=5, seed=52, rmin=0.15, rmax=0.3)
    vel = inputs.fourier_velocity(g, seed=53)
    dt = inputs.cfl_dt(g, *vel)
    for btype in ((1, 1, 1, 1, 1, 1), (0, 2, 2, 1, 2, 0)):
        blk = api.Block(g, btype=btype, big_phi=0.3)
        ob = O.Block(g, btype=btype, big_phi=0.3)
        ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
        p = api.ad_RKTVD3(blk, dev(phi), ucn, dt, mode="exact")
        p = api.ad_RKTVD3(blk, p, ucn, dt, mode="exact")
        ref = O.rk3(ob, phi, O.contra(ob, *vel), dt, nsteps=2)
        assert np.array_equal(host(p), ref), (shape, btype, linf_rel(host(p), ref))
        pf = api.ad_RKTVD3(blk, dev(phi), ucn, dt, mode="fast")
        assert linf_rel(host(pf), O.rk3(ob, phi, O.contra(ob, *vel), dt)) <= 1e-12


@pytest.mark.parametrize("n", [24, 65])
def test_rk3_folded_vs_oracle(n):
    """The frozen-velocity (folded) form stays within the fast-arithmetic
    bound of the reference, with and without an advection mask."""
    g, phi, vel, dt = _case(n, 61)
    blk = api.Block(g)
    ob = O.Block(g)
    ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
    V = api.fold_velocity(blk, ucn, dt)
    U = O.contra(ob, *vel)
    p = dev(phi)
    for _ in range(3):
        p = api.ad_RKTVD3_folded(blk, p, V)
    assert linf_rel(host(p), O.rk3(ob, phi, U, dt, nsteps=3)) <= 3e-12
    rng = np.random.default_rng(62)
    mask = (rng.uniform(size=g.shape) > 0.2).astype(np.int32)
    pm = api.ad_RKTVD3_folded(blk, dev(phi), V, adv_mask=dev(mask, torch.int32))
    assert linf_rel(host(pm), O.rk3(ob, phi, U, dt, mask=mask)) <= 1e-12


def test_fold_rejects_nonpositive_jacobian():
    g = inputs.uniform_grid(8, 8, 8, h=0.1)
    g.aj = -g.aj
    blk = api.Block(g)
    ucn = [dev(np.ones(g.shape)) for _ in range(3)]
    with pytest.raises(api.LsmError):
        api.fold_velocity(blk, ucn, 1e-3)


BC_SHAPES = [(20, 12, 9), (83, 11, 14), (35, 8, 20), (48, 4, 6)]
BC_TYPES = [(1, 1, 1, 1, 1, 1), (0, 2, 1, 2, 1, 0), (2, 2, 0, 0, 0, 1), (2, 0, 2, 1, 0, 0)]


@pytest.mark.parametrize("shape", BC_SHAPES)
def test_rk3_boundary_types_and_masks(shape):
    """Every boundary-type mix, with and without an advection mask: exact
    arithmetic is bit-identical to the oracle over two steps; fast and folded
    arithmetic stay within the fast-arithmetic bound."""
    ni, nj, nk = shape
    g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=71, nj=nj, nk=nk)
    phi = inputs.spheres_sdf(g, nspheres=5, seed=72, rmin=0.15, rmax=0.3)
    vel = inputs.fourier_velocity(g, seed=73)
    dt = inputs.cfl_dt(g, *vel)
    rng = np.random.default_rng(74)
    mask = (rng.uniform(size=g.shape) > 0.25).astype(np.int32)
    for btype in BC_TYPES:
        blk = api.Block(g, btype=btype, big_phi=0.3)
        ob = O.Block(g, btype=btype, big_phi=0.3)
        U = O.contra(ob, *vel)
        ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
        V = api.fold_velocity(blk, ucn, dt)
        for m in (None, mask):
            md = None if m is None else dev(m, torch.int32)
            p = api.ad_RKTVD3(blk, dev(phi), ucn, dt, adv_mask=md, mode="exact")
            p = api.ad_RKTVD3(blk, p, ucn, dt, adv_mask=md, mode="exact")
            ref = O.rk3(ob, phi, U, dt, mask=m, nsteps=2)
            assert np.array_equal(host(p), ref), (shape, btype, m is None, linf_rel(host(p), ref))
            ref1 = O.rk3(ob, phi, U, dt, mask=m)
            pf = host(api.ad_RKTVD3(blk, dev(phi), ucn, dt, adv_mask=md, mode="fast"))
            assert linf_rel(pf, ref1) <= 1e-12, (shape, btype, "fast")
            pv = host(api.ad_RKTVD3_folded(blk, dev(phi), V, adv_mask=md))
            assert linf_rel(pv, ref1) <= 1e-12, (shape, btype, "folded")
