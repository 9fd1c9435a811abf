"""ctypes front end of the CPU restatement in lsm_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker.  Arrays follow the
package's convention: float64, shape (nk, nj, ni), i-fastest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsm_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OBlock(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32),
                ("btype", C.c_int32 * 6),
                ("dc", C.c_double), ("de", C.c_double), ("dz", C.c_double),
                ("big_phi", C.c_double),
                ("csi", _dp * 3), ("eta", _dp * 3), ("zet", _dp * 3),
                ("aj", _dp), ("x", _dp), ("y", _dp), ("z", _dp),
                ("nblk", C.c_int32), ("blank", (C.c_int32 * 10) * 4)]


class OReinit(C.Structure):
    _fields_ = [("epslsm", C.c_double), ("eps_reinit", C.c_double),
                ("rn_timestep", C.c_double), ("rn_itermax", C.c_int32),
                ("order_bc", C.c_int32), ("eno_bc", C.c_int32), ("lambda_method", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        bp = C.POINTER(OBlock)
        rp = C.POINTER(OReinit)
        L.oracle_contra.argtypes = [bp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_weno3_rhs.argtypes = [bp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.oracle_bcond.argtypes = [bp, _dp]
        L.oracle_rk3.argtypes = [bp, _dp, _dp, _ip, _dp, _dp, _dp, C.c_double]
        L.oracle_rk3_order.argtypes = [bp, _dp, _dp, _ip, _dp, _dp, _dp, C.c_double, C.c_int]
        L.oracle_upwind_rhs.argtypes = [bp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.oracle_normgrad.argtypes = [bp, _dp, _dp, _dp, C.c_int, C.c_int]
        L.oracle_reinit_eno2.argtypes = [bp, rp, _dp, _dp]
        L.oracle_reinit_eno2.restype = C.c_int
        L.oracle_sign_ini.argtypes = [bp, _dp, _dp, _dp, C.c_double, C.c_int]
        L.oracle_rh_rn.argtypes = [bp, _dp, _dp, _dp]
        L.oracle_lambda_rn.argtypes = [bp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
        L.oracle_lambda_rn_simpson.argtypes = [bp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
        L.oracle_reinit_godunov.argtypes = [bp, rp, C.c_int, C.c_int, _dp, _dp]
        L.oracle_water_volume.argtypes = [bp, _dp, _dp, _ip]
        L.oracle_levelset_step.argtypes = [bp, rp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_levelset_step.restype = C.c_double
        L.oracle_levelset_step_h.argtypes = [bp, rp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_levelset_step_h.restype = C.c_double
        L.oracle_geometric_reinit.argtypes = [bp, _dp, _dp, _ip]
        L.oracle_geometric_reinit.restype = C.c_int
        L.oracle_gradient_eno2.argtypes = [bp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
        L.oracle_calc_h.argtypes = [bp, _dp, _dp, _dp]
        L.oracle_calc_h.restype = C.c_int
        L.oracle_free_surface.argtypes = [bp, _dp, _dp, _dp]
        L.oracle_free_surface.restype = C.c_int
        L.oracle_weno3.argtypes = [C.c_double] * 5
        L.oracle_weno3.restype = C.c_double
        L.oracle_eno2.argtypes = [C.c_double] * 6 + [C.c_int, C.c_int]
        L.oracle_eno2.restype = C.c_double
        L.oracle_biased.argtypes = [C.c_double] * 4 + [C.c_int] * 3
        L.oracle_biased.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Block:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, grid, btype=(1, 1, 1, 1, 1, 1), big_phi=5.0, blanking=()):
        """blanking: up to 4 boxes ((ia, ib, ja, jb, ka, kb), (bt1, bt2, bt3, bt4))
        with 0-based node indices (the reference's li_blk_* - 1) and blktype(1:4)."""
        self.keep = [_c(a) for a in (grid.csi[0], grid.csi[1], grid.csi[2],
                                     grid.eta[0], grid.eta[1], grid.eta[2],
                                     grid.zet[0], grid.zet[1], grid.zet[2],
                                     grid.aj, grid.x, grid.y, grid.z)]
        nk, nj, ni = grid.shape
        s = OBlock()
        s.ni, s.nj, s.nk = ni, nj, nk
        s.btype[:] = list(btype)
        s.dc, s.de, s.dz = grid.dc, grid.de, grid.dz
        s.big_phi = big_phi
        k = self.keep
        s.csi[:] = [_p(k[0]), _p(k[1]), _p(k[2])]
        s.eta[:] = [_p(k[3]), _p(k[4]), _p(k[5])]
        s.zet[:] = [_p(k[6]), _p(k[7]), _p(k[8])]
        s.aj, s.x, s.y, s.z = _p(k[9]), _p(k[10]), _p(k[11]), _p(k[12])
        if len(blanking) > 4:
            raise ValueError("at most 4 blanking boxes")
        s.nblk = len(blanking)
        for q, (box, bt) in enumerate(blanking):
            s.blank[q][:] = [int(v) for v in box] + [int(v) for v in bt]
        self.s = s
        self.shape = grid.shape

    @property
    def ref(self):
        return C.byref(self.s)


def _reinit(epslsm, eps_reinit, rn_timestep, rn_itermax, order_bc=1, eno_bc=False):
    r = OReinit()
    r.epslsm, r.eps_reinit, r.rn_timestep = epslsm, eps_reinit, rn_timestep
    r.rn_itermax, r.order_bc, r.eno_bc = rn_itermax, order_bc, int(bool(eno_bc))
    return r


def contra(blk, u, v, w):
    u, v, w = _c(u), _c(v), _c(w)
    U = [np.empty(blk.shape) for _ in range(3)]
    lib().oracle_contra(blk.ref, _p(u), _p(v), _p(w), _p(U[0]), _p(U[1]), _p(U[2]))
    return U


def weno3_rhs(blk, phi, U, mask=None):
    phi = _c(phi)
    mask = np.ones(blk.shape, np.int32) if mask is None else _c(mask, np.int32)
    rh = np.zeros(blk.shape)
    Uc = [_c(a) for a in U]
    lib().oracle_weno3_rhs(blk.ref, _p(phi), _p(Uc[0]), _p(Uc[1]), _p(Uc[2]),
                           mask.ctypes.data_as(_ip), _p(rh))
    return rh


def bcond(blk, phi):
    out = _c(phi).copy()
    lib().oracle_bcond(blk.ref, _p(out))
    return out


def rk3(blk, phi_n, U, dt, mask=None, nsteps=1, numiter=1, order_ad=1):
    """nsteps physical steps of numiter ad_RKTVD3 sub-steps each (order_ad 1:
    WENO3, 0: first-order upwind)."""
    mask = np.ones(blk.shape, np.int32) if mask is None else _c(mask, np.int32)
    Uc = [_c(a) for a in U]
    cur = _c(phi_n).copy()
    out = np.empty(blk.shape)
    for _ in range(nsteps * numiter):
        lib().oracle_rk3_order(blk.ref, _p(out), _p(cur), mask.ctypes.data_as(_ip),
                               _p(Uc[0]), _p(Uc[1]), _p(Uc[2]), float(dt), int(order_ad))
        cur, out = out, cur
    return cur


def upwind_rhs(blk, phi, U, mask=None, init=None):
    """calc_RH_AD (orderAD = 0); entries with mask <= 0 keep `init` (zeros)."""
    mask = np.ones(blk.shape, np.int32) if mask is None else _c(mask, np.int32)
    phi = _c(phi)
    Uc = [_c(a) for a in U]
    rh = np.zeros(blk.shape) if init is None else _c(init).copy()
    lib().oracle_upwind_rhs(blk.ref, _p(phi), _p(Uc[0]), _p(Uc[1]), _p(Uc[2]), mask.ctypes.data_as(_ip), _p(rh))
    return rh


def normgrad(blk, d, phi0, order_bc=1, eno_bc=False):
    d, phi0 = _c(d), _c(phi0)
    ng = np.zeros(blk.shape)
    lib().oracle_normgrad(blk.ref, _p(d), _p(phi0), _p(ng), int(order_bc), int(bool(eno_bc)))
    return ng


def gradient_eno2(blk, f, init, nc=3, order_bc=1, eno_bc=False):
    """calc_phi_gradient_ENO2 (nc = 3) / calc_h_gradient_ENO2 (nc = 2) of f;
    `init` (nc arrays) is what the output holds where the routine does not
    write (the eight vertices)."""
    f = _c(f)
    out = [_c(a).copy() for a in init]
    ptr = [_p(a) for a in out] + [None] * (3 - nc)
    lib().oracle_gradient_eno2(blk.ref, _p(f), ptr[0], ptr[1], ptr[2], int(nc), int(order_bc),
                               int(bool(eno_bc)))
    return out


def calc_h(blk, z, phi):
    """calc_h: (h, number of columns without a free surface)."""
    z, phi = _c(z), _c(phi)
    h = np.empty(blk.shape)
    missing = lib().oracle_calc_h(blk.ref, _p(z), _p(phi), _p(h))
    return h, missing


def free_surface(blk, z, phi):
    """z_free_surface of every column, shape (nj, ni); (zfs, missing)."""
    z, phi = _c(z), _c(phi)
    zfs = np.empty(blk.shape[1:])
    missing = lib().oracle_free_surface(blk.ref, _p(z), _p(phi), _p(zfs))
    return zfs, missing


def reinit_eno2(blk, phi0, epslsm, eps_reinit, rn_timestep, rn_itermax, order_bc=1, eno_bc=False):
    phi0 = _c(phi0)
    out = np.empty(blk.shape)
    r = _reinit(epslsm, eps_reinit, rn_timestep, rn_itermax, order_bc, eno_bc)
    it = lib().oracle_reinit_eno2(blk.ref, C.byref(r), _p(phi0), _p(out))
    return out, it


def sign_ini(blk, phi0, epslsm, m_sphi0=1):
    phi0 = _c(phi0)
    s = np.empty(blk.shape)
    g = np.empty(blk.shape)
    lib().oracle_sign_ini(blk.ref, _p(phi0), _p(s), _p(g), float(epslsm), int(m_sphi0))
    return s, g


def rh_rn(blk, d, sgn):
    d, sgn = _c(d), _c(sgn)
    rh = np.empty(blk.shape)
    lib().oracle_rh_rn(blk.ref, _p(d), _p(sgn), _p(rh))
    return rh


def lambda_rn(blk, phi0, ng0, d, dt, epslsm, method=0):
    """calc_lambda_RN: method 0 the pointwise quotient, 1 the filtered one
    (sussman_correction_method)."""
    phi0, ng0, d = _c(phi0), _c(ng0), _c(d)
    lam = np.empty(blk.shape)
    f = lib().oracle_lambda_rn_simpson if method == 1 else lib().oracle_lambda_rn
    f(blk.ref, _p(phi0), _p(ng0), _p(d), _p(lam), float(dt), float(epslsm))
    return lam


def reinit_godunov(blk, phi0, epslsm, rn_timestep, rn_itermax, m_sphi0=1, m_ti=1, lambda_method=0):
    phi0 = _c(phi0)
    out = np.empty(blk.shape)
    r = _reinit(epslsm, 0.0, rn_timestep, rn_itermax)
    r.lambda_method = int(lambda_method)
    lib().oracle_reinit_godunov(blk.ref, C.byref(r), int(m_sphi0), int(m_ti), _p(phi0), _p(out))
    return out


def water_volume(blk, phi):
    phi = _c(phi)
    vol = np.zeros(3)
    mask = np.zeros(blk.shape, np.int32)
    lib().oracle_water_volume(blk.ref, _p(phi), _p(vol), mask.ctypes.data_as(_ip))
    return vol, mask


def geometric_reinit(blk, phi0, pc):
    """geometric_reinitialisation (no global mass correction): returns
    (phicorrected, ids) from phicorrected's initial values pc; ids = 1 at the
    interface nodes, -1 at their neighbours, 0 elsewhere."""
    phi0 = _c(phi0)
    out = _c(pc).copy()
    ids = np.zeros(blk.shape, np.int32)
    lib().oracle_geometric_reinit(blk.ref, _p(phi0), _p(out), ids.ctypes.data_as(_ip))
    return out, ids


def levelset_step(blk, phi, u, v, w, delti_lsm, epslsm, eps_reinit, rn_timestep,
                  rn_itermax, order_bc=1, eno_bc=False, numiter_lsm=1,
                  call_levelset=True, call_reinit=True, rn_freq=1, ntime=1, hybrid=False):
    """One call of levelsetmethod (non-hybrid); returns (phi_new, volume)."""
    out = _c(phi).copy()
    u, v, w = _c(u), _c(v), _c(w)
    r = _reinit(epslsm, eps_reinit, rn_timestep, rn_itermax, order_bc, eno_bc)
    vol = lib().oracle_levelset_step_h(blk.ref, C.byref(r), _p(out), _p(u), _p(v), _p(w),
                                       int(numiter_lsm), float(delti_lsm), int(bool(call_levelset)),
                                       int(bool(call_reinit)), int(rn_freq), int(ntime), int(bool(hybrid)))
    return out, vol
