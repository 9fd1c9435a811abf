"""Batched multi-block advection (lsm_batch_*).

Parity rests on a property of the reference's block decomposition: with the
halo exchanged after every RK stage (ad_RKTVD3.F90:83-84) and the forced
WENO3 stencils and bcond_lsm applied only at physical faces, every owned node
of a block sees exactly the values the single-block step sees at the same
global node, in the same operation order.  So a connected batch must
reproduce, bit for bit, the single-block library step (itself bit-identical to
the oracle) on the assembled grid, in every launch form and tile shape; an
independent batch must reproduce each instance's own single-block step.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from lsm_amd import _lib, batch as B, inputs  # noqa: E402


def _case(lattice=(3, 2, 2), nmin=6, nmax=14, ghost=3, seed=11, btype=(1, 1, 1, 1, 1, 1), align=16):
    lay, grids, phi, vel, amps = B.multiblock_case(lattice=lattice, nmin=nmin, nmax=nmax, ghost=ghost, seed=seed,
                                                   nspheres=6, btype=btype, align=align)
    return lay, grids, phi, vel, amps


def test_multiblock_grid_is_one_grid():
    """Per-block coordinates and metrics at owned nodes equal those of the
    assembled single-block grid (same mapping, differences inside the domain)."""
    lay, grids, phi, vel, amps = _case()
    N = [sum(s) for s in lay.sizes]
    x, y, z = B.block_coordinates(lay, amps, [0, 0, 0], N)
    G = inputs.compute_metrics(inputs.Grid(x, y, z))
    assert np.array_equal(B.assemble(lay, [g.x for g in grids]), x)
    assert np.array_equal(B.assemble(lay, [g.aj for g in grids]), G.aj)
    for c in range(3):
        for name in ("csi", "eta", "zet"):
            assert np.array_equal(B.assemble(lay, [getattr(g, name)[c] for g in grids]), getattr(G, name)[c])
    # per-block arrays carry the ghost layers; rows padded to 16 doubles, the
    # first owned column of every row on a 128-byte line, arrays disjoint
    end = 0
    for b in range(len(lay.blocks)):
        assert grids[b].x.shape == lay.shape(b)
        assert lay.pitch[b] % 16 == 0 and lay.pitch[b] >= lay.shape(b)[2]
        assert (lay.offsets[b] + lay.ghost) % 16 == 0
        assert lay.offsets[b] >= end
        end = lay.offsets[b] + int(np.prod(lay.alloc_shape(b)))
    assert end <= lay.total
    dense = B.lattice_layout(lay.sizes, ghost=lay.ghost, align=0)
    for b in range(len(dense.blocks)):
        assert dense.offsets[b] % 32 == 0 and dense.pitch[b] == dense.shape(b)[2]


def _desc(lay, blocks=None, ghost=None, connected=None, lattice=None):
    arr = (_lib.LsmBatchBlock * len(lay.blocks))()
    for b, s in enumerate(lay.blocks):
        e = arr[b]
        e.ni, e.nj, e.nk = s.ni, s.nj, s.nk
        e.lattice[:] = list(s.lattice)
        e.phys_faces = s.phys_faces
        e.btype[:] = list(s.btype)
        e.big_phi = 5.0
        e.dc = e.de = e.dz = 1.0
        e.offset = lay.offsets[b]
        e.pitch = lay.pitch[b]
    if blocks:
        blocks(arr)
    d = _lib.LsmBatchDesc()
    d.nblocks = len(lay.blocks)
    d.ghost = lay.ghost if ghost is None else ghost
    d.connected = (1 if lay.connected else 0) if connected is None else connected
    d.lattice[:] = list(lay.lattice if lattice is None else lattice)
    d.blocks = C.cast(arr, C.POINTER(_lib.LsmBatchBlock))
    d.aj = 0x1000   # never dereferenced by the validation
    d.total = lay.total
    return d, arr


@pytest.mark.parametrize("what", ["faces", "sizes", "lattice", "ghost", "overlap", "independent_ghost", "pitch"])
def test_batch_create_rejects_bad_layouts(what):
    """Descriptor validation (host logic, before any device call)."""
    L = _lib.load()
    lay = B.lattice_layout(((6, 7), (5, 6), (8, 5)), ghost=3)
    kw = {}
    if what == "faces":
        kw["blocks"] = lambda a: setattr(a[0], "phys_faces", 0x3F)
    elif what == "sizes":
        kw["blocks"] = lambda a: setattr(a[1], "nj", 9)
    elif what == "lattice":
        kw["lattice"] = (2, 2, 3)
    elif what == "ghost":
        kw["ghost"] = 1
    elif what == "overlap":
        kw["blocks"] = lambda a: setattr(a[3], "offset", lay.total - 10)
    elif what == "independent_ghost":
        kw["connected"] = 0
    elif what == "pitch":
        kw["blocks"] = lambda a: setattr(a[2], "pitch", a[2].ni + 5)   # below ni + 2 ghost
    d, keep = _desc(lay, **kw)
    h = C.c_void_p()
    rc = L.lsm_batch_create(C.byref(d), C.byref(h))
    assert rc == -1, (rc, L.lsm_last_error())
    assert not h.value


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.gpu
@pytest.mark.parametrize("ghost", [2, 3])
def test_halo_fill_gives_owner_values(ghost):
    """Every ghost node inside the domain receives the value its owner block
    holds for that global node (edges and corners included); ghost nodes
    outside the domain are left alone."""
    lay, grids, _, _, _ = _case(lattice=(3, 3, 2), ghost=ghost, seed=12)
    bt = B.Batch(lay, grids)
    N = [sum(s) for s in lay.sizes]
    fields = []
    for b, s in enumerate(lay.blocks):
        shp = lay.shape(b)
        f = np.full(shp, -7.0)
        Gk, Gj, Gi = np.meshgrid(*[np.arange(shp[a]) - ghost + s.gstart[2 - a] for a in range(3)], indexing="ij")
        code = Gi + 1000.0 * Gj + 1e6 * Gk
        f[ghost:ghost + s.nk, ghost:ghost + s.nj, ghost:ghost + s.ni] = code[ghost:ghost + s.nk,
                                                                          ghost:ghost + s.nj, ghost:ghost + s.ni]
        fields.append((f, code, (Gi >= 0) & (Gi < N[0]) & (Gj >= 0) & (Gj < N[1]) & (Gk >= 0) & (Gk < N[2])))
    t = bt.pack([f for f, _, _ in fields])
    bt.halo_fill(t)
    got = bt.unpack(t)
    for (f, code, inside), out in zip(fields, got):
        assert np.array_equal(out[inside], code[inside])
        assert np.all(out[~inside] == -7.0)


def _single_block_steps(lay, grids, phi, vel, dt, nsteps, mode, btype):
    """The single-block library step on the assembled grid."""
    from lsm_amd import api
    N = [sum(s) for s in lay.sizes]
    amps = None  # noqa: F841
    g = inputs.Grid(*[B.assemble(lay, [getattr(gd, c) for gd in grids]) for c in "xyz"])
    g.aj = B.assemble(lay, [gd.aj for gd in grids])
    for name in ("csi", "eta", "zet"):
        setattr(g, name, np.stack([B.assemble(lay, [getattr(gd, name)[c] for gd in grids]) for c in range(3)]))
    blk = api.Block(g, btype=btype)
    ucn = api.rhs_contra_j(blk, *[_dev(B.assemble(lay, [v[c] for v in vel])) for c in range(3)])
    cur = _dev(B.assemble(lay, phi))
    for _ in range(nsteps):
        cur = api.ad_RKTVD3(blk, cur, ucn, dt, mode=mode)
    assert tuple(cur.shape) == (N[2], N[1], N[0])
    return cur.cpu().numpy()


# (form, tile rows, planes per tile, persistent workgroups per CU, tile columns)
CONFIGS = [("queue", 4, 8, 0, 64), ("queue", 8, 4, 2, 64), ("grid", 4, 16, 0, 64), ("grid", 8, 8, 0, 64),
           ("perblock", 4, 8, 0, 64), ("queue", 8, 8, 0, 32), ("grid", 4, 4, 0, 32)]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("btype", [(1, 1, 1, 1, 1, 1), (0, 2, 1, 2, 1, 0)])
@pytest.mark.parametrize("align", [16, 0])
def test_connected_batch_matches_single_block(mode, btype, align):
    lay, grids, phi, vel, _ = _case(btype=btype, align=align)
    dt = B.cfl_dt(lay, grids, vel)
    ref = _single_block_steps(lay, grids, phi, vel, dt, 3, mode, btype)
    bt = B.Batch(lay, grids)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    for form, ty, kc, wg, w in CONFIGS:
        bt.set_config(form, ty, kc, wg, tile_cols=w)
        assert bt.get_config() == {"form": form, "tile_cols": w, "tile_rows": ty, "kchunk": kc, "wg_per_cu": wg}
        cur = bt.halo_fill(bt.pack(phi))
        nxt = bt.empty()
        for _ in range(3):
            bt.rk3_step(cur, ucn, dt, out=nxt, mode=mode)
            cur, nxt = nxt, cur
        got = B.assemble(lay, bt.unpack(cur))
        assert np.array_equal(got, ref), (form, ty, kc, float(np.max(np.abs(got - ref))))


@pytest.mark.gpu
@pytest.mark.parametrize("align", [16, 0])
def test_independent_batch_matches_each_instance(align):
    from lsm_amd import api
    shapes = [(12, 9, 10), (17, 6, 8), (9, 14, 7), (33, 5, 6), (8, 8, 21)]
    lay = B.independent_layout(shapes, align=align)
    grids, phis, vels = [], [], []
    for q, (ni, nj, nk) in enumerate(shapes):
        gd = inputs.sinusoidal_cube_grid(ni, amp=0.03, seed=40 + q, nj=nj, nk=nk)
        grids.append(gd)
        phis.append(inputs.spheres_sdf(gd, nspheres=3, seed=50 + q, rmin=0.2, rmax=0.3))
        vels.append(inputs.fourier_velocity(gd, seed=60 + q))
    dt = min(inputs.cfl_dt(gd, *v) for gd, v in zip(grids, vels))
    bt = B.Batch(lay, grids)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vels]) for c in range(3)])
    refs = []
    for gd, p, v in zip(grids, phis, vels):
        blk = api.Block(gd)
        u = api.rhs_contra_j(blk, *[_dev(a) for a in v])
        refs.append(api.ad_RKTVD3(blk, api.ad_RKTVD3(blk, _dev(p), u, dt), u, dt).cpu().numpy())
    for form, ty, kc, wg, w in CONFIGS:
        bt.set_config(form, ty, kc, wg, tile_cols=w)
        a = bt.rk3_step(bt.pack(phis), ucn, dt)
        out = bt.unpack(bt.rk3_step(a, ucn, dt))
        for r, o in zip(refs, out):
            assert np.array_equal(o, r), form


@pytest.mark.gpu
def test_graph_replays_and_autotune():
    """A captured step replayed several times, with eager work and
    synchronisations between the replays, gives the eager step every time;
    the autotuner leaves a configuration that still reproduces it, and a
    second batch of the same shapes takes that configuration from the cache."""
    lay, grids, phi, vel, _ = _case(lattice=(2, 3, 2), seed=13)
    dt = B.cfl_dt(lay, grids, vel)
    bt = B.Batch(lay, grids)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    p0 = bt.halo_fill(bt.pack(phi))
    own = lambda t: B.assemble(lay, bt.unpack(t))  # noqa: E731  (ghosts beyond physical faces are not written)
    ref = own(bt.rk3_step(p0, ucn, dt))
    cfg, ms = bt.autotune(p0, ucn, dt, reps=2, force=True)
    assert ms is not None and ms > 0
    assert np.array_equal(own(bt.rk3_step(p0, ucn, dt)), ref), cfg
    out = bt.empty()
    work = bt.workspace()
    bt.graph_capture(out, p0, ucn, dt, work=work)
    for r in range(4):
        out.fill_(-1.0)
        bt.graph_launch()
        torch.cuda.synchronize()
        assert np.array_equal(own(out), ref), r
        bt.rk3_step(p0, ucn, dt)   # eager work between replays
        torch.cuda.synchronize()
    bt2 = B.Batch(lay, grids)
    cfg2, ms2 = bt2.autotune(p0, ucn, dt)
    assert cfg2 == cfg and ms2 is None


@pytest.mark.gpu
@pytest.mark.parametrize("reconf", ["reinit_tile", "tiles", "form", "autotune"])
def test_reconfiguration_drops_the_graph(reconf):
    """A captured step holds the device tables of its configuration; after a
    reconfiguration the batch refuses to replay it (the launch fails on the
    host, no kernel runs) until the step is captured again."""
    from lsm_amd._lib import LsmError
    lay, grids, phi, vel, _ = _case(lattice=(2, 2, 1), seed=16)
    dt = B.cfl_dt(lay, grids, vel)
    bt = B.Batch(lay, grids)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    p0 = bt.halo_fill(bt.pack(phi))
    bt.set_config("grid", 4, 8, 0)
    out = bt.empty()
    bt.graph_capture(out, p0, ucn, dt)
    bt.graph_launch()
    torch.cuda.synchronize()
    if reconf == "reinit_tile":
        bt.set_reinit_tile(32)
    elif reconf == "tiles":
        bt.set_config("grid", 8, 8, 0)
    elif reconf == "form":
        bt.set_config("queue", 4, 8, 0)
    else:
        bt.autotune(p0, ucn, dt, reps=1, force=True)
    with pytest.raises(LsmError):
        bt.graph_launch()
    ref = B.assemble(lay, bt.unpack(bt.rk3_step(p0, ucn, dt)))
    bt.graph_capture(out, p0, ucn, dt)
    out.fill_(0.0)
    bt.graph_launch()
    torch.cuda.synchronize()
    assert np.array_equal(B.assemble(lay, bt.unpack(out)), ref)


def _params(epslsm, dtau, nit=4, eps=1e-9, order=1, eno=False):
    from lsm_amd import api
    return api.ReinitParams(epslsm=epslsm, RNtimestep=dtau, RNitermax=nit, epsReinitialisation=eps,
                            OrderReinitialisationBoundaries=order, ENOBCReinitialisation=eno)


@pytest.mark.gpu
@pytest.mark.parametrize("form,rtw", [("grid", 64), ("grid", 32), ("grid", 16), ("perblock", 0)])
def test_independent_batch_reinit_matches_each_instance(form, rtw):
    """Each instance's reinitialisation (own exit test: the planar instance
    passes it at once and runs no iteration) equals the single-block call,
    for every tile width of the batched stage kernel."""
    from lsm_amd import api
    shapes = [(12, 9, 10), (17, 6, 8), (9, 14, 7), (12, 10, 9)]
    lay = B.independent_layout(shapes)
    grids, phis = [], []
    for q, (ni, nj, nk) in enumerate(shapes):
        gd = inputs.sinusoidal_cube_grid(ni, amp=0.02 if q == 3 else 0.04, seed=70 + q, nj=nj, nk=nk)
        grids.append(gd)
        if q == 3:
            phis.append((gd.z - 0.45 - 0.1 * gd.x) * 1.05)
        else:
            phis.append(inputs.spheres_sdf(gd, nspheres=3, seed=80 + q, rmin=0.2, rmax=0.35) * (0.5 + (gd.x - 0.3) ** 2))
    hmin = min(inputs.spacing_bounds(gd)[0] for gd in grids)
    hmax = max(inputs.spacing_bounds(gd)[1] for gd in grids)
    prm = _params(1.5 * hmax, 0.5 * hmin, nit=4, eps=0.1, order=2)
    bt = B.Batch(lay, grids)
    bt.set_config(form)
    bt.set_reinit_tile(rtw)
    print(f"[batch] independent reinit form={form} tile={bt.get_reinit_tile()}", flush=True)
    out, iters = bt.reinit(bt.pack(phis), prm)
    got = bt.unpack(out)
    its = iters.cpu().numpy()
    assert its[3] == 0 and all(its[:3] == 4), its
    for q, (gd, p) in enumerate(zip(grids, phis)):
        blk = api.Block(gd)
        ref, it = api.reinitialisation_test2(blk, _dev(p), prm)
        assert int(it.cpu()[0]) == its[q]
        assert np.array_equal(got[q], ref.cpu().numpy()), q


@pytest.mark.gpu
@pytest.mark.parametrize("ghost,order,eno,rtw", [(3, 1, False, 0), (2, 2, True, 32), (3, 3, False, 16),
                                                 (3, 2, False, 64)])
def test_connected_batch_reinit_matches_blockwise_reference(ghost, order, eno, rtw):
    """The reinitialisation of a connected batch is the reference's per-rank
    loop (ghost layers frozen, ENO2 through the ghosts at internal faces,
    one-sided formulas at physical faces), checked against the FP64 PyTorch
    re-expression of it block by block (1e-13: device sin/cos)."""
    import torch_ref as T
    lay, grids, phi, vel, _ = _case(ghost=ghost, seed=14)
    phi = [p * (0.5 + (gd.x - 0.3) ** 2) for p, gd in zip(phi, grids)]
    hmin = min(inputs.spacing_bounds(gd)[0] for gd in grids)
    hmax = max(inputs.spacing_bounds(gd)[1] for gd in grids)
    prm = _params(1.5 * hmax, 0.5 * hmin, nit=3, order=order, eno=eno)
    bt = B.Batch(lay, grids)
    bt.set_reinit_tile(rtw)
    print(f"[batch] connected reinit ghost={ghost} tile={bt.get_reinit_tile()}", flush=True)
    out, iters = bt.reinit(bt.halo_fill(bt.pack(phi)), prm)
    assert np.all(iters.cpu().numpy() == 3)
    got = bt.unpack(out)
    phi_h = bt.unpack(bt.halo_fill(bt.pack(phi)))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)  # noqa: E731
    for b, (gd, s) in enumerate(zip(grids, lay.blocks)):
        phys = tuple(bool(s.phys_faces >> q & 1) for q in range(6))
        met = ([t(c) for c in gd.csi], [t(c) for c in gd.eta], [t(c) for c in gd.zet])
        ref, it = T.reinit(t(phi_h[b]), met, (1.0, 1.0, 1.0), prm.epslsm, 1e-9, prm.RNtimestep, 3, order, eno,
                           ghost=ghost, phys=phys)
        assert it == 3
        r = ref.numpy()
        assert np.max(np.abs(got[b] - r)) <= 1e-13 * np.max(np.abs(r)), b
        g = ghost
        assert np.array_equal(got[b][:g], phi_h[b][:g])   # ghost layers untouched


@pytest.mark.gpu
def test_connected_batch_bcond_and_step_composition():
    """bcond_lsm of a connected batch equals the single-block bcond_lsm on the
    assembled grid; the full step equals rk3 -> reinit -> bcond -> halo fill;
    a graph of the full step replays it."""
    from lsm_amd import api
    btype = (1, 0, 1, 2, 1, 1)
    lay, grids, phi, vel, _ = _case(btype=btype, seed=15)
    bt = B.Batch(lay, grids)
    f = bt.halo_fill(bt.pack(phi))
    got = B.assemble(lay, bt.unpack(bt.bcond(f.clone())))
    g = inputs.Grid(*[B.assemble(lay, [getattr(gd, c) for gd in grids]) for c in "xyz"])
    g.aj = B.assemble(lay, [gd.aj for gd in grids])
    for name in ("csi", "eta", "zet"):
        setattr(g, name, np.stack([B.assemble(lay, [getattr(gd, name)[c] for gd in grids]) for c in range(3)]))
    blk = api.Block(g, btype=btype)
    ref = api.bcond_lsm(blk, _dev(B.assemble(lay, phi))).cpu().numpy()
    assert np.array_equal(got, ref)
    dt = B.cfl_dt(lay, grids, vel)
    hmin = min(inputs.spacing_bounds(gd)[0] for gd in grids)
    hmax = max(inputs.spacing_bounds(gd)[1] for gd in grids)
    prm = _params(1.5 * hmax, 0.5 * hmin, nit=4)
    ucn = bt.contra_velocity(*[bt.pack([v[c] for v in vel]) for c in range(3)])
    step = bt.levelset_step(f, ucn, dt, prm)
    x = bt.rk3_step(f, ucn, dt)
    y, _ = bt.reinit(x, prm)
    comp = bt.halo_fill(bt.bcond(y))
    own = lambda t: B.assemble(lay, bt.unpack(t))  # noqa: E731
    assert np.array_equal(own(step), own(comp))
    out = bt.empty()
    bt.graph_capture(out, f, ucn, dt, params=prm)
    for r in range(3):
        out.fill_(0.0)
        bt.graph_launch()
        torch.cuda.synchronize()
        assert np.array_equal(own(out), own(comp)), r
        bt.levelset_step(f, ucn, dt, prm)
        torch.cuda.synchronize()
