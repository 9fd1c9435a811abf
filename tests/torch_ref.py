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
