"""Geometric (hybrid) reinitialisation: geometric_reinitialisation.F90 with
GetTetrahedraList, GetIsosurfaceAndDistances and NeighboursPhiCorrection
(GlobalMassCorrection = .false., the reference's default), and the hybrid
path of levelsetmethod (advection restricted to the AdvectionNodes band,
+-9.9 away from the interface).

CPU: the oracle against golden vectors of the compiled reference (bit for
bit) and against what the geometry fixes (distances to a plane through grid
nodes, to a tilted plane, to a sphere).  GPU: the HIP entry points against
the oracle, bit for bit."""
import numpy as np
import pytest

import oracle as O
from fixtures import load
from lsm_amd import inputs

EPS = 2.220446049250313e-16


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    diff = int(np.sum(a != b))
    assert diff == 0, f"{diff} of {a.size} values differ, max |d| = {np.max(np.abs(a - b))}"


def _initial(phi, base=None):
    """phicorrected before geometric_reinitialisation (levelsetmethod.F90:300-301)."""
    pc = np.zeros_like(phi) if base is None else base.copy()
    pc[phi < -EPS] = -9.9
    pc[phi > EPS] = 9.9
    return pc


CASES = [("hybrid_cube", "geometric"), ("hybrid_cube", "levelset"), ("hybrid_channel", "geometric"),
         ("hybrid_channel", "levelset")] + [("hybrid_planes", f"geometric_{p}") for p in ("flat", "tilt", "diag", "sphere")]


def _run_oracle(fx, key):
    blk = O.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    phi = fx.inp("plane_" + key.split("_", 1)[1]) if fx.name == "hybrid_planes" else fx.inp("phi")
    cl = fx.kw(key, "call_levelset")
    if cl is None:
        cl = bool(fx.d["p_call_levelset"]) if "p_call_levelset" in fx.d else True
    out = phi.copy()
    for s in range(fx.kw(key, "nsteps", 1)):
        out, _ = O.levelset_step(blk, out, *fx.velocity, fx.par("delti_lsm"), fx.par("epslsm"),
                                 fx.par("eps_reinit"), fx.par("rn_timestep"), 4, call_levelset=cl, hybrid=True,
                                 ntime=1 + s)
    return out


@pytest.mark.parametrize("name,key", CASES, ids=[f"{a}-{b}" for a, b in CASES])
def test_oracle_hybrid_levelsetmethod_golden(name, key):
    fx = load(name)
    _eq(_run_oracle(fx, key), fx.out(key))


def test_oracle_geometric_distance_to_a_node_plane():
    """phi = z - c with c on a node layer: the interface nodes are the layer and
    its two neighbour layers; every node within 4 layers gets |z - c| exactly
    (to rounding) with the sign of phi, the others keep +-9.9."""
    g = inputs.uniform_grid(10, 9, 14, h=0.125)
    c = 6 * 0.125
    phi = g.z - c
    pc, ids = O.geometric_reinit(O.Block(g), phi, _initial(phi))
    dz = np.abs(g.z - c)
    band = dz <= 5 * 0.125 + 1e-12
    np.testing.assert_allclose(np.abs(pc[band]), dz[band], rtol=0, atol=1e-14)
    assert np.all(np.sign(pc[band & (phi != 0)]) == np.sign(phi[band & (phi != 0)]))
    assert np.all(np.abs(pc[~band]) == 9.9)
    assert set(np.unique(ids)) <= {-1, 0, 1}
    assert np.all(ids[np.abs(dz) < 1e-12] == 1)


def test_oracle_geometric_distance_to_a_tilted_plane():
    """The zero set of a linear phi is a plane and the piecewise-linear
    isosurface is exact.  Every candidate point of the distance search lies on
    that plane -- triangle vertices, points of edges, and the 'projection' of
    CheckProjectionPointOnATriangle3, which weights P1, P2, P3 with the
    barycentric weights of P3, P2 and P1 (an affine combination, so a point of
    the plane, but not the foot of the perpendicular) -- so the distances are
    never below the exact plane distance and, away from the domain faces and
    within two spacings of the plane, exceed it by a small fraction of h.  The
    offset keeps every node at least 1e-3 away from the plane: an edge of the
    isosurface shorter than rounding (two intersection points at a node where
    phi is zero) makes DistancePointLine divide by its length, and the
    reference's value there is not a distance."""
    g = inputs.uniform_grid(16, 15, 14, h=0.1)
    nrm = np.array([0.3, 0.2, -1.0])
    phi = nrm[0] * g.x + nrm[1] * g.y + nrm[2] * g.z + 0.6537
    assert np.min(np.abs(phi)) > 1e-3
    pc, ids = O.geometric_reinit(O.Block(g), phi, _initial(phi))
    dist = phi / np.linalg.norm(nrm)
    near = ids != 0
    assert np.all(np.abs(pc[near]) >= np.abs(dist[near]) - 1e-12)
    assert np.all(np.sign(pc[near & (np.abs(phi) > 1e-12)]) == np.sign(phi[near & (np.abs(phi) > 1e-12)]))
    inner = np.zeros(g.shape, bool)
    inner[5:-5, 5:-5, 5:-5] = True
    sel = inner & near & (np.abs(dist) <= 0.2)
    assert sel.sum() > 10
    assert np.max(np.abs(pc[sel]) - np.abs(dist[sel])) <= 0.01


def test_oracle_geometric_sphere_is_second_order():
    """Distance to a sphere: the isosurface is piecewise linear, so the error of
    the interface-node distances falls like h^2."""
    errs = []
    for n in (17, 33):
        h = 1.0 / (n - 1)
        g = inputs.uniform_grid(n, n, n, h=h)
        r = np.sqrt((g.x - 0.5) ** 2 + (g.y - 0.47) ** 2 + (g.z - 0.52) ** 2)
        phi = (0.3 - r) * (1.0 + 0.5 * (g.x - 0.5))        # same zero set, not a distance
        pc, ids = O.geometric_reinit(O.Block(g), phi, _initial(phi))
        sel = ids == 1
        errs.append(np.max(np.abs(pc[sel] - (0.3 - r[sel]))))
    assert errs[1] < errs[0] / 3.0, errs


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("name,key", CASES, ids=[f"{a}-{b}" for a, b in CASES])
def test_gpu_hybrid_levelsetmethod_golden(name, key):
    """levelsetmethod with hybrid_reinitialisation on the GPU (reference
    advection arithmetic) equals the compiled reference bit for bit; the
    pre-advection volumes to 1e-13."""
    from lsm_amd import api
    print(f"[geometric] hybrid levelsetmethod {name}/{key}", flush=True)
    fx = load(name)
    blk = api.Block(fx.grid, btype=fx.par("btype"), big_phi=fx.par("big_phi"))
    phi = _dev(fx.inp("plane_" + key.split("_", 1)[1]) if name == "hybrid_planes" else fx.inp("phi"))
    cl = fx.kw(key, "call_levelset")
    if cl is None:
        cl = bool(fx.d["p_call_levelset"]) if "p_call_levelset" in fx.d else True
    prm = api.ReinitParams(epslsm=fx.par("epslsm"), RNtimestep=fx.par("rn_timestep"), RNitermax=4,
                           epsReinitialisation=fx.par("eps_reinit"), hybrid_reinitialisation=True)
    u, v, w = (_dev(a) for a in fx.velocity)
    vols = []
    for s in range(fx.kw(key, "nsteps", 1)):
        vol = api.levelsetmethod(blk, phi, phi.clone(), u, v, w, prm, fx.par("delti_lsm"),
                                 call_levelsetmethod=cl, ntime=1 + s, mode="exact")
        vols.append(float(_host(vol).sum()))
    _eq(_host(phi), fx.out(key))
    gv = fx.d[f"out_{key}_volume"]
    assert np.max(np.abs(np.array(vols) - gv)) <= 1e-13 * np.max(gv)


def _seeded(kind, shape):
    ni, nj, nk = shape
    if kind == "spheres":
        g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=21, nj=nj, nk=nk)
        phi = inputs.spheres_sdf(g, nspheres=6, seed=22, rmin=0.08, rmax=0.25)
        phi = phi * (0.7 + (g.y - 0.3) ** 2)
    elif kind == "channel":
        g = inputs.channel_grid(ni, nj, nk)
        phi = inputs.solitary_phi(g, H=0.35)
    elif kind == "noise":                          # every node an interface node, slivers
        g = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=24, nj=nj, nk=nk)
        phi = np.random.default_rng(25).uniform(-0.5, 0.5, g.shape)
    elif kind == "steps":                          # values in {-2..2}: exact zeros everywhere
        g = inputs.uniform_grid(ni, nj, nk, h=0.37)
        phi = np.random.default_rng(26).integers(-2, 3, g.shape).astype(np.float64)
    else:                                          # zero on node layers and diagonals
        g = inputs.uniform_grid(ni, nj, nk, h=1.0 / 32)
        phi = np.minimum(g.z - 10.0 / 32, g.x - g.y + 3.0 / 32)
    return g, phi


GEO = [("spheres", (67, 45, 38), 1.0), ("spheres", (64, 64, 64), 1.0), ("channel", (150, 9, 41), 1.0),
       ("nodes", (40, 33, 29), 1.0), ("noise", (23, 19, 17), 1.0), ("steps", (21, 20, 19), 1.0),
       ("spheres", (31, 29, 27), 1.3)]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,shape,scale", GEO, ids=[f"{k}-{s[0]}x{s[1]}x{s[2]}-{c}" for k, s, c in GEO])
def test_gpu_geometric_matches_oracle(kind, shape, scale):
    """geometric_reinitialisation alone on larger blocks: phicorrected and
    InterfaceNodesID equal the oracle bit for bit (the isosurface store grows
    and is reused across these calls).  The distance search skips triangles
    by error-bounded lower bounds; noise (sliver triangles, every node an
    interface node), exact zeros everywhere and |phicorrected| above 10
    (scale 1.3: skips that are valid only below 10 must not happen) are the
    cases where a loose bound would show."""
    import torch
    from lsm_amd import api
    print(f"[geometric] oracle parity {kind} {shape} scale {scale}", flush=True)
    g, phi = _seeded(kind, shape)
    pc0 = _initial(phi, base=phi) * scale
    ref, ref_ids = O.geometric_reinit(O.Block(g), phi, pc0)
    blk = api.Block(g)
    pc = _dev(pc0)
    ids = blk.empty(torch.int32)
    api.geometric_reinitialisation(blk, _dev(phi), pc, ids)
    _eq(_host(ids), ref_ids)
    _eq(_host(pc), ref)
    assert int(np.sum(ref_ids == 1)) > 0


@pytest.mark.gpu
def test_gpu_hybrid_step_matches_oracle():
    """Three hybrid levelsetmethod steps (band-restricted WENO3 advection,
    geometric reinitialisation, bcond_lsm) on a seeded curvilinear block with
    mixed boundary types equal the oracle bit for bit."""
    from lsm_amd import api
    print("[geometric] hybrid steps vs oracle", flush=True)
    g, phi = _seeded("spheres", (53, 41, 37))
    u, v, w = inputs.fourier_velocity(g, seed=23)
    dt = inputs.cfl_dt(g, u, v, w)
    bt = (1, 0, 2, 1, 1, 2)
    oblk = O.Block(g, btype=bt, big_phi=0.4)
    ref = phi.copy()
    for s in range(3):
        ref, _ = O.levelset_step(oblk, ref, u, v, w, dt, 0.05, 1e-5, 1e-3, 4, hybrid=True, ntime=1 + s)
    blk = api.Block(g, btype=bt, big_phi=0.4)
    prm = api.ReinitParams(epslsm=0.05, RNtimestep=1e-3, RNitermax=4, epsReinitialisation=1e-5,
                           hybrid_reinitialisation=True)
    p = _dev(phi)
    du, dv, dw = (_dev(a) for a in (u, v, w))
    for s in range(3):
        api.levelsetmethod(blk, p, p.clone(), du, dv, dw, prm, dt, ntime=1 + s, mode="exact",
                           want_volume=False)
    _eq(_host(p), ref)


@pytest.mark.gpu
def test_gpu_hybrid_rejected_where_unsupported():
    """The step adjoint and the batched form cover the Sussman path only."""
    import torch
    from lsm_amd import api
    from lsm_amd.api import LsmError
    g, phi = _seeded("spheres", (20, 18, 16))
    blk = api.Block(g)
    prm = api.ReinitParams(epslsm=0.05, RNtimestep=1e-3, hybrid_reinitialisation=True)
    U = api.rhs_contra_j(blk, *[_dev(a) for a in inputs.fourier_velocity(g, seed=1)])
    with pytest.raises(LsmError, match="hybrid"):
        api.levelset_step_adjoint(blk, _dev(phi), U, 1e-3, prm, _dev(np.ones(g.shape)))
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_concurrent_streams_do_not_share_library_memory():
    """Two geometric reinitialisations (isosurface store kept by the library)
    and two Sussman reinitialisations with a NULL workspace (library scratch),
    issued on two streams of default priority at once, give the results of
    the same calls run one after the other, bit for bit."""
    import ctypes as C
    import torch
    from lsm_amd import _lib, api
    L = _lib.load()
    cases = []
    for q, n in enumerate((40, 46)):
        g = inputs.sinusoidal_cube_grid(n, amp=0.03, seed=300 + q, nj=n - 3, nk=n - 6)
        phi = inputs.spheres_sdf(g, nspheres=5, seed=310 + q, rmin=0.1, rmax=0.25)
        hmin, hmax = inputs.spacing_bounds(g)
        cases.append((api.Block(g), phi, api.ReinitParams(epslsm=1.5 * hmax, RNtimestep=0.5 * hmin, RNitermax=4)))

    def run(blk, phi, prm, stream):
        p0 = torch.as_tensor(phi, device="cuda")
        pc = torch.full_like(p0, 9.9) * torch.sign(p0)
        ids = torch.zeros(blk.shape, dtype=torch.int32, device="cuda")
        out = blk.empty()
        it = torch.zeros(1, dtype=torch.int32, device="cuda")
        s = C.c_void_p(stream.cuda_stream)
        ws = blk.workspace(_lib.LSM_OP_REINIT_GEOMETRIC)
        torch.cuda.current_stream().synchronize()   # inputs written on torch's stream
        _lib.check(L.lsm_reinit_geometric(blk.ref, C.c_void_p(p0.data_ptr()), C.c_void_p(pc.data_ptr()),
                                          C.c_void_p(ids.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel(), s),
                   "lsm_reinit_geometric")
        c = prm.c()
        _lib.check(L.lsm_reinit_sussman_eno2(blk.ref, C.byref(c), C.c_void_p(p0.data_ptr()),
                                             C.c_void_p(out.data_ptr()), C.c_void_p(it.data_ptr()), None, 0, s),
                   "lsm_reinit_sussman_eno2")
        return pc, ids, out, it, ws, p0

    s0 = torch.cuda.Stream()
    s1 = torch.cuda.Stream()
    ser = []
    for (blk, phi, prm) in cases:
        r = run(blk, phi, prm, s0)
        s0.synchronize()
        ser.append([t.cpu() for t in r[:4]])
    torch.cuda.synchronize()
    for rep in range(3):
        print(f"[geometric] two streams, repetition {rep}", flush=True)
        r0 = run(*cases[0], s0)
        r1 = run(*cases[1], s1)
        torch.cuda.synchronize()
        for got, ref in ((r0, ser[0]), (r1, ser[1])):
            for a, b in zip(got[:4], ref):
                assert torch.equal(a.cpu(), b), rep
