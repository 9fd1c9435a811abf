"""The hand-written adjoint of ad_RKTVD3 (lsm_rk3_step_adjoint, lsm_amd.autograd).

Checked three ways:
  * an FP64 PyTorch re-expression of the forward (tests/torch_ref.py, CPU) is
    first pinned to the oracle, then its autograd gradients are compared with
    the HIP adjoint (tolerance 1e-10 relative: same derivatives, different
    summation order);
  * central finite differences of the oracle forward (bit-identical to the
    compiled reference Fortran) along seeded random directions, for phi^n and
    for U/J (tolerance 1e-6 relative, the target of the port);
  * checkpointed trajectories give the gradients of the step-by-step chain.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402
import torch_ref as T  # noqa: E402
from lsm_amd import inputs  # noqa: E402

CASES = [((12, 10, 9), (1, 1, 1, 1, 1, 1), False), ((11, 9, 8), (0, 2, 1, 2, 1, 0), True),
         ((9, 12, 10), (2, 0, 2, 1, 0, 1), True)]


def _case(shape, seed=81):
    ni, nj, nk = shape
    g = inputs.sinusoidal_cube_grid(ni, amp=0.04, seed=seed, nj=nj, nk=nk)
    phi = inputs.spheres_sdf(g, nspheres=3, seed=seed + 1, rmin=0.2, rmax=0.35)
    vel = inputs.fourier_velocity(g, seed=seed + 2)
    dt = inputs.cfl_dt(g, *vel)
    return g, phi, vel, dt


def _cpu(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)


def _ref_grad(g, phi, U, dt, btype, big, mask, w, nsteps):
    p = _cpu(phi).requires_grad_(True)
    Ut = [_cpu(u).requires_grad_(True) for u in U]
    out = p
    m = None if mask is None else _cpu(mask)
    for _ in range(nsteps):
        out = T.rk3_step(out, Ut, _cpu(g.aj), dt, (g.dc, g.de, g.dz), btype, big, m)
    J = (out * _cpu(w)).sum()
    J.backward()
    return out.detach().numpy(), p.grad.numpy(), [u.grad.numpy() for u in Ut]


@pytest.mark.parametrize("shape,btype,with_mask", CASES)
def test_torch_reexpression_matches_oracle(shape, btype, with_mask):
    g, phi, vel, dt = _case(shape)
    ob = O.Block(g, btype=btype, big_phi=0.3)
    U = O.contra(ob, *vel)
    mask = (np.random.default_rng(3).uniform(size=g.shape) > 0.3).astype(np.int32) if with_mask else None
    ref = O.rk3(ob, phi, U, dt, mask=mask, nsteps=2)
    out = _cpu(phi)
    for _ in range(2):
        out = T.rk3_step(out, [_cpu(u) for u in U], _cpu(g.aj), dt, (g.dc, g.de, g.dz), btype, 0.3,
                         None if mask is None else _cpu(mask))
    d = np.max(np.abs(out.numpy() - ref)) / np.max(np.abs(ref))
    assert d <= 1e-15, d


def _fd(fun, x, e, h):
    return (fun(x + h * e) - fun(x - h * e)) / (2.0 * h)


def _u_directions(rng, U):
    """Random U/J directions that leave out nodes where |U/J| < 1e-4: there a
    step of 1e-6 could flip the upwind choice, where the step is not
    differentiable and a central difference measures the jump."""
    return [rng.standard_normal(u.shape) * (np.abs(u) > 1e-4) for u in U]


def test_torch_reexpression_gradient_matches_fd_of_oracle():
    """Pins the re-expression's autograd gradient: central differences of the
    oracle forward along random directions."""
    g, phi, vel, dt = _case((10, 9, 8), seed=91)
    btype = (1, 2, 0, 1, 1, 0)
    ob = O.Block(g, btype=btype, big_phi=5.0)
    U = O.contra(ob, *vel)
    w = np.random.default_rng(5).standard_normal(g.shape)
    _, gphi, gU = _ref_grad(g, phi, U, dt, btype, 5.0, None, w, 2)
    rng = np.random.default_rng(6)
    for _ in range(4):
        e = rng.standard_normal(g.shape)
        fd = _fd(lambda x: np.sum(w * O.rk3(ob, x, U, dt, nsteps=2)), phi, e, 1e-6)
        an = np.sum(gphi * e)
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, (fd, an)
    e = _u_directions(rng, U)
    fdu = _fd(lambda s: np.sum(w * O.rk3(ob, phi, [U[c] + s * e[c] for c in range(3)], dt, nsteps=2)),
              0.0, 1.0, 1e-7)
    anu = sum(np.sum(gU[c] * e[c]) for c in range(3))
    assert abs(fdu - anu) <= 1e-6 * abs(anu) + 1e-9, (fdu, anu)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,btype,with_mask", CASES)
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_adjoint_matches_autograd(shape, btype, with_mask, mode):
    from lsm_amd import api, autograd
    g, phi, vel, dt = _case(shape)
    ob = O.Block(g, btype=btype, big_phi=0.3)
    U = O.contra(ob, *vel)
    rng = np.random.default_rng(7)
    mask = (rng.uniform(size=g.shape) > 0.3).astype(np.int32) if with_mask else None
    w = rng.standard_normal(g.shape)
    _, gphi, gU = _ref_grad(g, phi, U, dt, btype, 0.3, mask, w, 2)
    blk = api.Block(g, btype=btype, big_phi=0.3)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    p = dev(phi).requires_grad_(True)
    Ud = [dev(u).requires_grad_(True) for u in U]
    md = None if mask is None else dev(mask).to(torch.int32)
    out = p
    for _ in range(2):
        out = autograd.rk3_step(blk, out, Ud, dt, mode=mode, adv_mask=md)
    (out * dev(w)).sum().backward()
    tol = 1e-10
    rel = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))  # noqa: E731
    assert rel(p.grad.cpu().numpy(), gphi) <= tol, rel(p.grad.cpu().numpy(), gphi)
    for c in range(3):
        assert rel(Ud[c].grad.cpu().numpy(), gU[c]) <= tol, (c, rel(Ud[c].grad.cpu().numpy(), gU[c]))


@pytest.mark.gpu
def test_adjoint_matches_fd_of_oracle():
    """Directional derivatives of J = sum(w * phi_T) after 3 steps: the HIP
    adjoint against central differences of the oracle forward (1e-6)."""
    from lsm_amd import api, autograd
    g, phi, vel, dt = _case((16, 14, 12), seed=101)
    btype = (1, 1, 0, 2, 1, 1)
    ob = O.Block(g, btype=btype, big_phi=5.0)
    U = O.contra(ob, *vel)
    w = np.random.default_rng(8).standard_normal(g.shape)
    blk = api.Block(g, btype=btype, big_phi=5.0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    p = dev(phi).requires_grad_(True)
    Ud = [dev(u).requires_grad_(True) for u in U]
    out = autograd.advect(blk, p, Ud, dt, 3, mode="exact", checkpoint_every=2)
    (out * dev(w)).sum().backward()
    gphi = p.grad.cpu().numpy()
    gU = [u.grad.cpu().numpy() for u in Ud]
    rng = np.random.default_rng(9)
    for _ in range(8):
        e = rng.standard_normal(g.shape)
        fd = _fd(lambda x: np.sum(w * O.rk3(ob, x, U, dt, nsteps=3)), phi, e, 1e-6)
        an = float(np.sum(gphi * e))
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, (fd, an)
    for _ in range(4):
        # U/J is O(10) here: a step of 1e-7 keeps the central difference's
        # truncation error (which falls as h^2: 1.4e-6 relative at 1e-6 on this
        # case) well below the 1e-6 bound
        e = _u_directions(rng, U)
        fdu = _fd(lambda s: np.sum(w * O.rk3(ob, phi, [U[c] + s * e[c] for c in range(3)], dt, nsteps=3)),
                  0.0, 1.0, 1e-7)
        anu = float(sum(np.sum(gU[c] * e[c]) for c in range(3)))
        assert abs(fdu - anu) <= 1e-6 * abs(anu) + 1e-9, (fdu, anu)


@pytest.mark.gpu
def test_checkpointed_trajectory_matches_step_chain():
    """advect() with checkpoints every 1, 2 and 5 steps gives the gradients of
    the step-by-step chain (dJ/dphi0 bit for bit), and the velocity gradient
    reaches the Cartesian velocity through contravariant()."""
    from lsm_amd import api, autograd
    g, phi, vel, dt = _case((20, 16, 14), seed=111)
    blk = api.Block(g)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    w = dev(np.random.default_rng(10).standard_normal(g.shape))
    grads = []
    for every in (None, 1, 2, 5):
        p = dev(phi).requires_grad_(True)
        uvw = [dev(a).requires_grad_(True) for a in vel]
        Ud = autograd.contravariant(blk, *uvw)
        if every is None:
            out = p
            for _ in range(5):
                out = autograd.rk3_step(blk, out, Ud, dt)
        else:
            out = autograd.advect(blk, p, Ud, dt, 5, checkpoint_every=every)
        (out * w).sum().backward()
        grads.append([p.grad.cpu().numpy()] + [a.grad.cpu().numpy() for a in uvw])
    for gset in grads[1:]:
        # dJ/dphi0: the same adjoint operations in the same order
        assert np.array_equal(gset[0], grads[0][0])
        # dJ/du: per-step contributions summed in another order (in place in the
        # trajectory, by autograd across steps in the chain)
        for a, b in zip(gset[1:], grads[0][1:]):
            assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # contravariant() equals the kernel's rhs_contra_j
    Uk = api.rhs_contra_j(blk, *(dev(a) for a in vel))
    Ut = autograd.contravariant(blk, *(dev(a) for a in vel))
    for a, b in zip(Uk, Ut):
        assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,btype,with_mask,rows", [
    ((70, 33, 20), (1, 1, 1, 1, 1, 1), False, "16"), ((70, 33, 20), (0, 2, 1, 2, 1, 0), True, "16"),
    ((61, 14, 37), (2, 0, 2, 1, 0, 1), True, "8"), ((9, 12, 10), (1, 2, 0, 1, 1, 0), False, "16")])
def test_fused_adjoint_matches_two_kernel_form(shape, btype, with_mask, rows, monkeypatch):
    """The fused adjoint stage kernels (partials kept on chip, several tiles in
    i, j and k: the first fused form and the lean one) against the two-kernel
    form with the partials in HBM."""
    from lsm_amd import api
    g, phi, vel, dt = _case(shape, seed=121)
    blk = api.Block(g, btype=btype, big_phi=0.3)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    rng = np.random.default_rng(12)
    md = dev((rng.uniform(size=g.shape) > 0.3).astype(np.int32)) if with_mask else None
    ucn = api.rhs_contra_j(blk, *(dev(a) for a in vel))
    lam = dev(rng.standard_normal(g.shape))
    res = {}
    for form in (rows, "2", "0"):
        monkeypatch.setenv("LSM_ADJ_FUSED", form)
        aU = [blk.zeros() for _ in range(3)]
        adj = api.ad_RKTVD3_adjoint(blk, dev(phi), ucn, dt, lam, adv_mask=md, mode="exact", adj_ucn=aU)
        res[form] = [t.cpu().numpy() for t in [adj] + aU]
    for form in (rows, "2"):
        for a, b in zip(res[form], res["0"]):
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), (form, np.max(np.abs(a - b)) / np.max(np.abs(b)))
