"""An FP64 PyTorch re-expression of ad_RKTVD3 (WENO3 right-hand side, SSP-RK3
stages, bcond_lsm) for checking the hand-written adjoint with autograd on
small grids.  Test infrastructure: operation order follows the reference
(ad_RKTVD3.F90:35-86, calc_RH_AD_WENO3.F90:195-254, AdvectionMethods.F90:
195-245, bcond_lsm.F90:53-287) so the forward matches the oracle to rounding.
Arrays are shaped (nk, nj, ni); axis 2 is i.
"""
import torch

C = [(1.0, 0.0, 1.0), (3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)]


def weno3(f0, f1, f2, f3, ws):
    pos = ws > 0.0
    fL = torch.where(pos, f0, f3)
    fC = torch.where(pos, f1, f2)
    fR = torch.where(pos, f2, f1)
    d0, d1, e = 2.0 / 3.0, 1.0 / 3.0, 0.000001
    b0 = (fC - fR) * (fC - fR)
    b1 = (fL - fC) * (fL - fC)
    a0 = d0 / ((b0 + e) * (b0 + e))
    a1 = d1 / ((b1 + e) * (b1 + e))
    w0 = a0 / (a0 + a1)
    w1 = a1 / (a0 + a1)
    u0 = 0.5 * fC + 0.5 * fR
    u1 = (-0.5) * fL + 1.5 * fC
    return u0 * w0 + u1 * w1


def _shift(a, axis, o):
    # value at index + o along axis (wrapping; wrapped entries are never used)
    return torch.roll(a, shifts=-o, dims=axis)


def rhs(phi, U, dd):
    """calc_RH_AD_WENO3 on every node (only interior entries are meaningful)."""
    shape = phi.shape
    out = None
    for d, axis in ((0, 2), (1, 1), (2, 0)):
        n = shape[axis]
        idx = torch.arange(n, device=phi.device).view([-1 if q == axis else 1 for q in range(3)])
        fm = (idx == 1).expand(shape)
        fp = (idx == n - 2).expand(shape)
        m2, m1 = _shift(phi, axis, -2), _shift(phi, axis, -1)
        p1, p2 = _shift(phi, axis, 1), _shift(phi, axis, 2)
        ws = U[d]
        zero = torch.zeros_like(phi)
        fplus = weno3(m1, phi, p1, torch.where(fp, zero, p2), torch.where(fp, torch.ones_like(ws), ws))
        fminus = weno3(torch.where(fm, zero, m2), m1, phi, p1, torch.where(fm, -torch.ones_like(ws), ws))
        L = ws * (dd[d] * (fplus - fminus))
        out = L if out is None else out + L
    return out


def _bc(bt, p1, p2, self_, big, is_im):
    if bt == 0:
        return p1
    if bt == 1:
        a2, b2, a1, b1 = 4.0 / 3.0, -1.0 / 3.0, 1.0, 0.0
        ex = (1.0 + torch.copysign(torch.ones_like(p2), torch.abs(p2) - big)) / 2.0
        ex = ex.detach()
        ta = a2 + ex * (a1 - a2)
        tb = b2 + ex * (b1 - b2)
        return ta * p1 + tb * p2
    if bt == 2 and is_im:
        return 2.0 * p1 - 1.0 * p2
    return self_


def bcond(phi, btype, big):
    """bcond_lsm on all six faces (physical), in the order i1, im, j1, jm, k1, km."""
    phi = phi.clone()
    for axis, (blo, bhi) in ((2, btype[0:2]), (1, btype[2:4]), (0, btype[4:6])):
        n = phi.shape[axis]

        def at(q):
            sl = [slice(None)] * 3
            sl[axis] = q
            return tuple(sl)
        if blo != 2:
            phi[at(0)] = _bc(blo, phi[at(1)], phi[at(2)], phi[at(0)], big, False)
        phi[at(n - 1)] = _bc(bhi, phi[at(n - 2)], phi[at(n - 3)], phi[at(n - 1)], big, axis == 2)
    return phi


def rk3_step(phi_n, U, aj, dt, dd, btype, big, mask=None):
    shape = phi_n.shape
    inner = torch.zeros(shape, dtype=torch.bool, device=phi_n.device)
    inner[1:-1, 1:-1, 1:-1] = True
    upd = torch.ones(shape, dtype=torch.bool, device=phi_n.device) if mask is None else mask > 0
    prev = phi_n
    for c1, c2, c3 in C:
        rh = rhs(prev, U, dd)
        lin = c1 * phi_n + c2 * prev
        interior = lin - ((c3 * aj) * dt) * rh
        boundary = lin - ((c3 * aj) * dt) * 0.0
        new = torch.where(inner, interior, boundary)
        new = torch.where(upd, new, prev)
        prev = bcond(new, btype, big)
    return prev


# ---------------------------------------------------------------------------
# Sussman reinitialisation (reinitialisation_test2.F90:42-210 with
# calc_NormGrad_d.F90:39-2446, AdvectionMethods.F90:64-302,576-587 and
# global_lsm.F90:120-133,200-219), FP64, every node at once.
# ---------------------------------------------------------------------------
EPS_SIMS = 2.220446049250313e-16
PI_REF = 3.141592653589793


def _sign1(x):
    return torch.copysign(torch.ones_like(x), x)


def eno2(f0, fLL, fL, fC, fR, fRR, lo, hi):
    a = (fR - 2.0 * fC) + fL
    bp = (fRR - 2.0 * fR) + fC
    dp = torch.where(hi & ~lo, (fR - fC) - 0.5 * a,
                     (fR - fC) - 0.5 * torch.where(torch.abs(bp) < torch.abs(a), bp, a))
    bm = (fC - 2.0 * fL) + fLL
    dm = torch.where(lo, (fC - fL) + 0.5 * a,
                     (fC - fL) + 0.5 * torch.where(torch.abs(bm) < torch.abs(a), bm, a))
    s = torch.where(torch.abs(f0) > EPS_SIMS, _sign1(f0), torch.zeros_like(f0))
    c1 = (s * (fR - fC) < 0.0) & (s * (fC - fL) < -(s * (fR - fC)))
    c2 = (s * (fC - fL) > 0.0) & (s * (fR - fC) > -(s * (fC - fL)))
    return torch.where(c1, dp, torch.where(c2, dm, 0.5 * (dp + dm)))


def _eno2_boundary(f0, f1, f2, as_):
    D1 = as_ * (f1 - f0)
    D2 = as_ * (f2 - f1)
    D3 = as_ * (f2 - f0) / 2.0
    a1, a2, a3 = torch.abs(D1), torch.abs(D2), torch.abs(D3)
    r = D1
    r = torch.where((a1 >= a3) & (a2 >= a3), D3, r)
    r = torch.where((a1 < a3) & (a2 < a3), ((3.0 * D1 + 6.0 * D2) - D3) / 7.0, r)
    r = torch.where(a1 == a2, (4.0 * D1 + D2) / 5.0, r)
    r = torch.where(a2 < a1, (D1 + 2.0 * D2) / 3.0, r)
    r = torch.where(a1 < a2, (2.0 * D1 + D2) / 3.0, r)
    return torch.where(torch.abs(D1 - D2) < EPS_SIMS, D1, r)


def biased(f0, f1, f2, f3, order, as_, enobc):
    if order == 1:
        return as_ * (-(1.0 * f0) + 1.0 * f1)
    if order == 2:
        if enobc:
            return _eno2_boundary(f0, f1, f2, as_)
        return as_ * ((-(3.0 * f0) + 4.0 * f1) - 1.0 * f2) / 2.0
    return as_ * (((-(11.0 * f0) + 18.0 * f1) - 9.0 * f2) + 2.0 * f3) / 6.0


def normgrad(d, phi0, metrics, dd, order_bc=1, eno_bc=False, ghost=0, phys=(True,) * 6):
    """calc_NormGrad_d on the owned nodes of one block (0 on ghost nodes):
    `ghost` layers on every face, phys = physical flags of the faces (i low,
    i high, j low, j high, k low, k high); one-sided formulas on physical
    faces, ENO2 elsewhere (ghost values in the stencil at internal faces)."""
    nk, nj, ni = d.shape
    dev = d.device
    g = ghost
    ii = torch.arange(ni, device=dev).view(1, 1, -1).expand(d.shape)
    jj = torch.arange(nj, device=dev).view(1, -1, 1).expand(d.shape)
    kk = torch.arange(nk, device=dev).view(-1, 1, 1).expand(d.shape)
    idx = (ii, jj, kk)
    n = (ni, nj, nk)
    onl = [(idx[a] == g) & phys[2 * a] for a in range(3)]
    onh = [(idx[a] == n[a] - 1 - g) & phys[2 * a + 1] for a in range(3)]
    on = [onl[a] | onh[a] for a in range(3)]
    last = [n[a] - 1 - g - (1 if phys[2 * a + 1] else 0) for a in range(3)]
    stale = phi0[last[2], last[1], last[0]]
    p0 = torch.where(on[0] & ~on[1] & ~on[2], stale.expand(d.shape), phi0)
    zero = torch.zeros_like(d)
    gr = []
    for a, axis in ((0, 2), (1, 1), (2, 0)):
        sh = lambda o: _shift(d, axis, o)  # noqa: E731
        lo = (idx[a] == g + 1) & phys[2 * a]
        hi = (idx[a] == n[a] - 2 - g) & phys[2 * a + 1]
        inner = dd[a] * eno2(p0, torch.where(lo, zero, sh(-2)), sh(-1), d, sh(1), torch.where(hi, zero, sh(2)),
                             lo, hi)
        f_lo = dd[a] * biased(d, sh(1), sh(2), sh(3), order_bc, 1.0, eno_bc)
        f_hi = dd[a] * biased(d, sh(-1), sh(-2), sh(-3), order_bc, -1.0, eno_bc)
        gv = torch.where(onl[a], f_lo, torch.where(onh[a], f_hi, inner))
        if a == 0:
            dc2 = 0.5 * dd[0]
            edge = dc2 * ((-(1.0 * sh(2)) + 4.0 * sh(1)) - 3.0 * d)
            gv = torch.where(onl[0] & onl[2] & ~on[1], edge, gv)
        gr.append(gv)
    csi, eta, zet = metrics
    gx = (csi[0] * gr[0] + eta[0] * gr[1]) + zet[0] * gr[2]
    gy = (csi[1] * gr[0] + eta[1] * gr[1]) + zet[1] * gr[2]
    gz = (csi[2] * gr[0] + eta[2] * gr[1]) + zet[2] * gr[2]
    ng = torch.sqrt((gx * gx + gy * gy) + gz * gz)
    if g == 0:
        return ng
    own = (ii >= g) & (ii <= ni - 1 - g) & (jj >= g) & (jj <= nj - 1 - g) & (kk >= g) & (kk <= nk - 1 - g)
    return torch.where(own, ng, torch.zeros_like(ng))


def sign_phi0(p, eps):
    return torch.where(torch.abs(p) > eps, _sign1(p), p / eps - (1.0 / PI_REF) * torch.sin((PI_REF * p) / eps))


def deltaf(v, eps):
    return torch.where(torch.abs(v) > eps, torch.zeros_like(v),
                       (1.0 / (2.0 * eps)) * (1.0 + torch.cos((PI_REF * v) / eps)))


def lambda_mass(p, s, g, eps):
    dl = deltaf(p, eps)
    num = -((dl * s) * (1.0 - g))
    den = (dl * dl) * g
    small = torch.abs(den) < EPS_SIMS
    return torch.where(small, torch.zeros_like(den), num / torch.where(small, torch.ones_like(den), den))


def reinit(phi0, metrics, dd, epslsm, eps_reinit, dtau, itermax, order_bc=1, eno_bc=False, ghost=0,
           phys=(True,) * 6):
    """reinitialisation_test2 (non-hybrid) on one block: returns (phi_out,
    iterations).  With ghost layers (one rank of a decomposition) only owned
    nodes are updated and the ghost values stay those of phi0."""
    s = sign_phi0(phi0, epslsm)
    out = phi0
    kw = dict(ghost=ghost, phys=phys)
    if ghost:
        nk, nj, ni = phi0.shape
        own = torch.zeros(phi0.shape, dtype=torch.bool, device=phi0.device)
        own[ghost:nk - ghost, ghost:nj - ghost, ghost:ni - ghost] = True
    ng = normgrad(out, phi0, metrics, dd, order_bc, eno_bc, **kw)
    it = 0
    while it < itermax:
        mx, mn = float(ng.detach().max()), float(ng.detach().min())
        if abs(mx - 1.0) < eps_reinit and abs(mn - 1.0) < eps_reinit:
            break
        dn = out
        upd = (normgrad(dn, phi0, metrics, dd, order_bc, eno_bc, **kw) - 1.0 > eps_reinit).detach()
        if ghost:
            upd = upd & own
        for c1, c2, c3 in C:
            ng = normgrad(out, phi0, metrics, dd, order_bc, eno_bc, **kw)
            rhs = s * (ng - 1.0)
            lam = lambda_mass(out, s, ng, epslsm)
            aux = deltaf(out, epslsm) * ng
            rhs = rhs - lam * aux
            out = torch.where(upd, (c1 * dn + c2 * out) - (c3 * dtau) * rhs, out)
        it += 1
    return out, it
