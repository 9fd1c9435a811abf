"""PyTorch-facing level-set transport operators.

Each function mirrors one routine of the reference level-set path (same name,
same argument meaning) and runs it on the GPU through the C ABI of
liblsm_hip.so.  Fields are float64 CUDA tensors of shape (nk, nj, ni),
contiguous (i fastest, the reference's Fortran order).  Calls are queued on
torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from contextlib import contextmanager
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import LsmBlock, LsmReinitParams, LsmError, check

# "upwind": orderAD = 0 (first-order upwind right-hand side, calc_RH_AD.F90); forward steps only
MODES = {"fast": _lib.LSM_ADVECT_FAST, "exact": _lib.LSM_ADVECT_EXACT, "upwind": _lib.LSM_ADVECT_UPWIND1}


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _f64(t: torch.Tensor, shape, name: str):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor")
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous float64 CUDA tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != block shape {tuple(shape)}")
    return C.c_void_p(t.data_ptr())


def _i32(t: Optional[torch.Tensor], shape, name: str):
    if t is None:
        return None
    if not t.is_cuda or t.dtype != torch.int32 or not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous int32 CUDA tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != block shape {tuple(shape)}")
    return C.c_void_p(t.data_ptr())


# tensors allocated by this module inside a multi-instance recording, kept
# alive until the recorded launches have been issued (lsm_multi_end)
_recording: list = []


def _keep(t):
    if _recording:
        _recording[-1].append(t)
    return t


class MultiRecorder:
    def __init__(self, L):
        self._L = L
        self.count = 0

    def instance(self):
        """Starts the next instance: the library calls that follow belong to it."""
        check(self._L.lsm_multi_instance(), "lsm_multi_instance")
        self.count += 1


@contextmanager
def multi():
    """Multi-instance execution (lsm_multi_begin .. lsm_multi_end): inside the
    block, call rec.instance() before each instance's library calls; those
    calls record their launches instead of issuing them, and at the end of
    the block every kernel runs as one launch over all instances, on the
    current stream.  Only library calls may touch the recorded outputs inside
    the block (anything else would run before the recorded kernels)."""
    L = _lib.load()
    check(L.lsm_multi_begin(), "lsm_multi_begin")
    keep: list = []
    _recording.append(keep)
    rec = MultiRecorder(L)
    try:
        yield rec
    except BaseException:
        L.lsm_multi_cancel()
        raise
    finally:
        _recording.pop()
    check(L.lsm_multi_end(_stream()), "lsm_multi_end")
    del keep


class Block:
    """One curvilinear block resident on the GPU: metrics, coordinates and the
    face/boundary description (the host-associated state of levelsetmethod)."""

    def __init__(self, grid, btype: Sequence[int] = (1, 1, 1, 1, 1, 1), big_phi: float = 5.0,
                 device="cuda", blanking=()):
        """blanking: up to four regions ((ia, ib, ja, jb, ka, kb), (blktype 1..4)) with
        0-based node indices of the numpy arrays (the reference's li_blk_* - 1)."""
        def dev(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)

        self.shape = tuple(grid.shape)
        nk, nj, ni = self.shape
        self.csi = [dev(grid.csi[c]) for c in range(3)]
        self.eta = [dev(grid.eta[c]) for c in range(3)]
        self.zet = [dev(grid.zet[c]) for c in range(3)]
        self.aj = dev(grid.aj)
        self.x, self.y, self.z = dev(grid.x), dev(grid.y), dev(grid.z)
        s = LsmBlock()
        s.il, s.iu, s.jl, s.ju, s.kl, s.ku = 1, ni, 1, nj, 1, nk
        s.igp = s.jgp = s.kgp = 0
        s.phys_faces = _lib.LSM_FACE_ALL
        s.btype[:] = [int(b) for b in btype]
        s.dc, s.de, s.dz = grid.dc, grid.de, grid.dz
        s.big_phi = float(big_phi)
        s.csi[:] = [t.data_ptr() for t in self.csi]
        s.eta[:] = [t.data_ptr() for t in self.eta]
        s.zet[:] = [t.data_ptr() for t in self.zet]
        s.aj = self.aj.data_ptr()
        s.x, s.y, s.z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
        if len(blanking) > _lib.LSM_MAX_BLANK:
            raise ValueError(f"at most {_lib.LSM_MAX_BLANK} blanking regions")
        s.nblk = len(blanking)
        for q, (box, bt) in enumerate(blanking):
            B = s.blank[q]
            B.ia, B.ib, B.ja, B.jb, B.ka, B.kb = [int(v) + 1 for v in box]   # array indices (il = 1)
            B.blktype[:] = [int(v) for v in bt]
        self.blanking = tuple(blanking)
        self.desc = s
        self.device = self.aj.device
        # the reinitialisation's metric tensor, computed once for the block's grid
        self.gten = torch.empty(6 * ni * nj * nk, dtype=torch.float64, device=device)
        s.gten = self.gten.data_ptr()
        check(_lib.load().lsm_metric_tensor(self.ref, C.c_void_p(self.gten.data_ptr()), _stream()),
              "lsm_metric_tensor")

    @property
    def ref(self):
        return C.byref(self.desc)

    def empty(self, dtype=torch.float64):
        return _keep(torch.empty(self.shape, dtype=dtype, device=self.device))

    def zeros(self, dtype=torch.float64):
        return _keep(torch.zeros(self.shape, dtype=dtype, device=self.device))

    def workspace(self, op: int) -> torch.Tensor:
        n = _lib.load().lsm_workspace_bytes(self.ref, op)
        if n == 0:
            raise LsmError("lsm_workspace_bytes failed: " + _lib.load().lsm_last_error().decode())
        return _keep(torch.empty(n, dtype=torch.uint8, device=self.device))

    def adjoint_workspace(self, op: int, rn_itermax: int) -> torch.Tensor:
        """Scratch of the adjoints that keep every reinitialisation iteration."""
        n = _lib.load().lsm_adjoint_workspace_bytes(self.ref, op, int(rn_itermax))
        if n == 0:
            raise LsmError(f"lsm_adjoint_workspace_bytes failed (rn_itermax = {rn_itermax}): "
                           + _lib.load().lsm_last_error().decode())
        return _keep(torch.empty(n, dtype=torch.uint8, device=self.device))


@dataclass
class ReinitParams:
    """Pseudo-time reinitialisation parameters (global_lsm.F90 names)."""
    epslsm: float
    RNtimestep: float
    RNitermax: int = 4
    epsReinitialisation: float = 1e-5
    OrderReinitialisationBoundaries: int = 1
    ENOBCReinitialisation: bool = False
    M_Sphi0: int = 1
    M_TI: int = 1
    hybrid_reinitialisation: bool = False   # levelsetmethod only: geometric reinitialisation
    sussman_correction_method: int = 0      # Godunov path: 0 pointwise lambda, 1 filtered (testfilter)

    def c(self) -> LsmReinitParams:
        p = LsmReinitParams()
        p.epslsm = self.epslsm
        p.eps_reinit = self.epsReinitialisation
        p.rn_timestep = self.RNtimestep
        p.rn_itermax = int(self.RNitermax)
        p.order_bc = int(self.OrderReinitialisationBoundaries)
        p.eno_bc = int(bool(self.ENOBCReinitialisation))
        p.m_sphi0 = int(self.M_Sphi0)
        p.m_ti = int(self.M_TI)
        p.hybrid = int(bool(self.hybrid_reinitialisation))
        p.lambda_method = int(self.sussman_correction_method)
        return p


def _ws(blk: Block, op: int, work: Optional[torch.Tensor]):
    if work is None:
        work = blk.workspace(op)
    return C.c_void_p(work.data_ptr()), work.numel(), work


# ---------------------------------------------------------------------------
# advection
# ---------------------------------------------------------------------------

def rhs_contra_j(blk: Block, u, v, w):
    """rhs_contra_j.F90:2-50: contravariant velocity divided by the Jacobian,
    returned as three tensors (ucn_j(1:3,...))."""
    L = _lib.load()
    out = [blk.empty() for _ in range(3)]
    check(L.lsm_contra_velocity(blk.ref, _f64(u, blk.shape, "u"), _f64(v, blk.shape, "v"),
                                _f64(w, blk.shape, "w"), 1, *[C.c_void_p(t.data_ptr()) for t in out],
                                _stream()), "lsm_contra_velocity")
    return out


def calc_RH_AD_WENO3(blk: Block, phi, ucn, adv_mask=None, rhs=None, mode="fast"):
    """calc_RH_AD_WENO3.F90:1-407: WENO3 advection right-hand side on interior
    nodes (rhs is zeroed first, as ad_RKTVD3.F90:55 does)."""
    L = _lib.load()
    if rhs is None:
        rhs = blk.zeros()
    else:
        rhs.zero_()
    check(L.lsm_weno3_rhs(blk.ref, _f64(phi, blk.shape, "phi"), *[_f64(t, blk.shape, "ucn") for t in ucn],
                          _i32(adv_mask, blk.shape, "adv_mask"), _f64(rhs, blk.shape, "rhs"),
                          MODES[mode], _stream()), "lsm_weno3_rhs")
    return rhs


def calc_RH_AD(blk: Block, phi, ucn, adv_mask=None, out=None):
    """calc_RH_AD.F90:1-273 (orderAD = 0): first-order upwind right-hand side
    at every node with AdvectionNodes > 0 (rhs is zeroed first, as
    ad_RKTVD3.F90:55 does, unless `out` is given: its other entries are kept)."""
    if out is None:
        out = blk.zeros()
    check(_lib.load().lsm_upwind_rhs(blk.ref, _f64(phi, blk.shape, "phi"),
                                     *[_f64(t, blk.shape, "ucn") for t in ucn],
                                     _i32(adv_mask, blk.shape, "adv_mask"), _f64(out, blk.shape, "out"),
                                     _stream()), "lsm_upwind_rhs")
    return out


def bcond_lsm(blk: Block, phi):
    """bcond_lsm.F90:2-568, in place."""
    check(_lib.load().lsm_bcond(blk.ref, _f64(phi, blk.shape, "phi"), _stream()), "lsm_bcond")
    return phi


def ad_RKTVD3(blk: Block, phi_n, ucn, delti_lsm: float, adv_mask=None, out=None, mode="fast",
              work=None):
    """ad_RKTVD3.F90:1-88: phi^(n+1) = SSP-RK3 step of phi^n with the WENO3
    right-hand side (mode "fast" / "exact") or the first-order upwind one
    (mode "upwind", orderAD = 0)."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    wp, wn, keep = _ws(blk, _lib.LSM_OP_RK3, work)
    check(L.lsm_rk3_step(blk.ref, _f64(out, blk.shape, "out"), _f64(phi_n, blk.shape, "phi_n"),
                         _i32(adv_mask, blk.shape, "adv_mask"),
                         *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                         wp, wn, _stream()), "lsm_rk3_step")
    return out


def ad_RKTVD3_adjoint(blk: Block, phi_n, ucn, delti_lsm: float, adj_np1, adv_mask=None,
                      mode="fast", adj_ucn=None, out=None, work=None):
    """Reverse mode of ad_RKTVD3: given adj_np1 = dJ/dphi^(n+1), returns
    dJ/dphi^n; when adj_ucn (three tensors) is given, dJ/d(U/J) of this step
    is ADDED to it.  The forward stages are recomputed from phi_n."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    wp, wn, keep = _ws(blk, _lib.LSM_OP_RK3_ADJOINT, work)
    au = [_f64(t, blk.shape, "adj_ucn") for t in adj_ucn] if adj_ucn is not None else [None] * 3
    check(L.lsm_rk3_step_adjoint(blk.ref, _f64(phi_n, blk.shape, "phi_n"),
                                 _i32(adv_mask, blk.shape, "adv_mask"),
                                 *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                 _f64(adj_np1, blk.shape, "adj_np1"), _f64(out, blk.shape, "out"),
                                 *au, wp, wn, _stream()), "lsm_rk3_step_adjoint")
    return out


def fold_velocity(blk: Block, ucn, delti_lsm: float):
    """Frozen-velocity operator V_d = aj*delti_lsm*d_d*ucn_d for ad_RKTVD3_folded
    (requires aj > 0; raises LsmError otherwise)."""
    L = _lib.load()
    V = [blk.empty() for _ in range(3)]
    check(L.lsm_advect_fold(blk.ref, *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm),
                            *[C.c_void_p(t.data_ptr()) for t in V], _stream()), "lsm_advect_fold")
    return V


def ad_RKTVD3_folded(blk: Block, phi_n, V, adv_mask=None, out=None, work=None):
    """ad_RKTVD3 with the fast arithmetic on a folded frozen-velocity operator
    (fold_velocity); the step size is the one V was folded with."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    wp, wn, keep = _ws(blk, _lib.LSM_OP_RK3, work)
    check(L.lsm_rk3_step_folded(blk.ref, _f64(out, blk.shape, "out"), _f64(phi_n, blk.shape, "phi_n"),
                                _i32(adv_mask, blk.shape, "adv_mask"),
                                *[_f64(t, blk.shape, "V") for t in V], wp, wn, _stream()),
          "lsm_rk3_step_folded")
    return out


# ---------------------------------------------------------------------------
# reinitialisation
# ---------------------------------------------------------------------------

def calc_NormGrad_d(blk: Block, d, phizero, order_bc: int = 1, eno_bc: bool = False, out=None):
    """calc_NormGrad_d.F90:1-2448: ENO2 |grad d| with biased boundary derivatives."""
    if out is None:
        out = blk.empty()
    check(_lib.load().lsm_normgrad_eno2(blk.ref, _f64(d, blk.shape, "d"), _f64(phizero, blk.shape, "phizero"),
                                        int(order_bc), int(bool(eno_bc)), _f64(out, blk.shape, "out"),
                                        _stream()), "lsm_normgrad_eno2")
    return out


def calc_phi_gradient_ENO2(blk: Block, phi, order_bc: int = 1, eno_bc: bool = False, out=None):
    """calc_phi_gradient_ENO2.F90:1-2025: (dphi/dx, dphi/dy, dphi/dz) with the
    calc_NormGrad_d stencils and the node's own phi as the ENO2 upwind sign.
    The eight vertices are not written: they keep what `out` holds (zeros
    when `out` is None)."""
    if out is None:
        out = [blk.zeros() for _ in range(3)]
    check(_lib.load().lsm_phi_gradient_eno2(blk.ref, _f64(phi, blk.shape, "phi"), int(order_bc),
                                            int(bool(eno_bc)), *[_f64(o, blk.shape, "out") for o in out],
                                            _stream()), "lsm_phi_gradient_eno2")
    return out


def calc_h_gradient_ENO2(blk: Block, h, order_bc: int = 1, eno_bc: bool = False, out=None):
    """calc_h_gradient_ENO2.F90:1-2018: (dh/dx, dh/dy), vertices as in
    calc_phi_gradient_ENO2."""
    if out is None:
        out = [blk.zeros() for _ in range(2)]
    check(_lib.load().lsm_h_gradient_eno2(blk.ref, _f64(h, blk.shape, "h"), int(order_bc), int(bool(eno_bc)),
                                          *[_f64(o, blk.shape, "out") for o in out], _stream()),
          "lsm_h_gradient_eno2")
    return out


def calc_h(blk: Block, z, phi, out=None):
    """calc_h.F90:1-394 with find_z_free_surface.F90: h = z_fs - z in every
    column (0 where the column has no free surface).  Returns (h, missing)
    with missing a one-element int32 CUDA tensor: the number of columns
    without a free surface."""
    if out is None:
        out = blk.empty()
    missing = torch.zeros(1, dtype=torch.int32, device=blk.device)
    check(_lib.load().lsm_calc_h(blk.ref, _f64(z, blk.shape, "z"), _f64(phi, blk.shape, "phi"),
                                 _f64(out, blk.shape, "out"), C.c_void_p(missing.data_ptr()), _stream()),
          "lsm_calc_h")
    return out, missing


def free_surface(blk: Block, z, phi, out=None):
    """find_z_free_surface over every column (as calc_h obtains z_free_surface):
    returns (zfs, missing), zfs of shape (nj, ni) (0 in a column without a
    free surface), missing a one-element int32 CUDA tensor."""
    nk, nj, ni = blk.shape
    if out is None:
        out = torch.empty((nj, ni), dtype=torch.float64, device=blk.device)
    missing = torch.zeros(1, dtype=torch.int32, device=blk.device)
    check(_lib.load().lsm_free_surface(blk.ref, _f64(z, blk.shape, "z"), _f64(phi, blk.shape, "phi"),
                                       _f64(out, (nj, ni), "out"), C.c_void_p(missing.data_ptr()), _stream()),
          "lsm_free_surface")
    return out, missing


def free_surface_adjoint(blk: Block, z, phi, adj_zfs, out=None):
    """Reverse mode of free_surface: (d zfs / d phi)^T adj_zfs at every node."""
    nk, nj, ni = blk.shape
    if out is None:
        out = blk.empty()
    check(_lib.load().lsm_free_surface_adjoint(blk.ref, _f64(z, blk.shape, "z"), _f64(phi, blk.shape, "phi"),
                                               _f64(adj_zfs, (nj, ni), "adj_zfs"), _f64(out, blk.shape, "out"),
                                               _stream()), "lsm_free_surface_adjoint")
    return out


def reinitialisation_test2(blk: Block, phi0, params: ReinitParams, out=None, work=None):
    """reinitialisation_test2.F90:1-218 (non-hybrid): returns (phi_reinitialised,
    iterations) with iterations a one-element int32 CUDA tensor."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    iters = _keep(torch.zeros(1, dtype=torch.int32, device=blk.device))
    wp, wn, keep = _ws(blk, _lib.LSM_OP_REINIT_ENO2, work)
    p = params.c()
    check(L.lsm_reinit_sussman_eno2(blk.ref, C.byref(p), _f64(phi0, blk.shape, "phi0"),
                                    _f64(out, blk.shape, "out"), C.c_void_p(iters.data_ptr()),
                                    wp, wn, _stream()), "lsm_reinit_sussman_eno2")
    return out, iters


def geometric_reinitialisation(blk: Block, phizero, phicorrected, InterfaceNodesID=None, work=None):
    """geometric_reinitialisation.F90:1-94 (GlobalMassCorrection = .false.,
    TriangulationCase 1): phicorrected (updated in place and returned) gets the
    signed distance to the marching-tetrahedra isosurface of phizero at the
    interface nodes and min(|phicorrected|, distance) with its own sign at the
    nodes within 4 nodes of them.  InterfaceNodesID (int32, optional) receives
    1 / -1 / 0.  Needs the block's coordinates."""
    L = _lib.load()
    wp, wn, keep = _ws(blk, _lib.LSM_OP_REINIT_GEOMETRIC, work)
    check(L.lsm_reinit_geometric(blk.ref, _f64(phizero, blk.shape, "phizero"),
                                 _f64(phicorrected, blk.shape, "phicorrected"),
                                 _i32(InterfaceNodesID, blk.shape, "InterfaceNodesID"), wp, wn, _stream()),
          "lsm_reinit_geometric")
    return phicorrected


def reinitialisation_test2_adjoint(blk: Block, phi0, params: ReinitParams, adj_out, out=None,
                                  phi_out=None, work=None):
    """Reverse mode of reinitialisation_test2: given adj_out = dJ/dphi_out,
    returns dJ/dphi0.  The iterations are recomputed from phi0 (phi_out, if
    given, receives their result)."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    if work is None:
        work = blk.adjoint_workspace(_lib.LSM_OP_REINIT_ENO2_ADJOINT, params.RNitermax)
    p = params.c()
    check(L.lsm_reinit_sussman_eno2_adjoint(blk.ref, C.byref(p), _f64(phi0, blk.shape, "phi0"),
                                            _f64(adj_out, blk.shape, "adj_out"), _f64(out, blk.shape, "out"),
                                            _f64(phi_out, blk.shape, "phi_out") if phi_out is not None else None,
                                            C.c_void_p(work.data_ptr()), work.numel(), _stream()),
          "lsm_reinit_sussman_eno2_adjoint")
    return out


def bcond_lsm_adjoint(blk: Block, phi_bc, adj):
    """Reverse mode of bcond_lsm, in place on adj (phi_bc: the field after bcond_lsm)."""
    check(_lib.load().lsm_bcond_adjoint(blk.ref, _f64(phi_bc, blk.shape, "phi"), _f64(adj, blk.shape, "adj"),
                                        _stream()), "lsm_bcond_adjoint")
    return adj


def calc_sign_ini(blk: Block, phi0, epslsm: float, M_Sphi0: int = 1):
    """calc_sign_ini.F90:1-286: (s(phi0), |grad phi0| Godunov)."""
    s, g = blk.empty(), blk.empty()
    check(_lib.load().lsm_sign_ini(blk.ref, _f64(phi0, blk.shape, "phi0"), float(epslsm), int(M_Sphi0),
                                   _f64(s, blk.shape, "sign"), _f64(g, blk.shape, "ng"), _stream()),
          "lsm_sign_ini")
    return s, g


def calc_RH_RN(blk: Block, d, sign):
    """calc_RH_RN.F90:1-244: s * (|grad d| - 1)."""
    rh = blk.empty()
    check(_lib.load().lsm_rh_rn(blk.ref, _f64(d, blk.shape, "d"), _f64(sign, blk.shape, "sign"),
                                _f64(rh, blk.shape, "rh"), _stream()), "lsm_rh_rn")
    return rh


def calc_lambda_RN(blk: Block, phi0, normgrad0, d, timestep: float, epslsm: float,
                   sussman_correction_method: int = 0, work=None):
    """calc_lambda_RN.F90:1-122: the point-constant lambda (method 0) or the
    one filtered by testfilter (method 1)."""
    lam = blk.empty()
    L = _lib.load()
    args = (blk.ref, _f64(phi0, blk.shape, "phi0"), _f64(normgrad0, blk.shape, "ng0"), _f64(d, blk.shape, "d"),
            float(timestep), float(epslsm), _f64(lam, blk.shape, "lambda"))
    if sussman_correction_method == 1:
        wp, wn, keep = _ws(blk, _lib.LSM_OP_LAMBDA_RN_SIMPSON, work)
        check(L.lsm_lambda_rn_simpson(*args, wp, wn, _stream()), "lsm_lambda_rn_simpson")
    elif sussman_correction_method == 0:
        check(L.lsm_lambda_rn(*args, _stream()), "lsm_lambda_rn")
    else:
        raise ValueError(f"sussman_correction_method = {sussman_correction_method} (0 or 1)")
    return lam


def reinitialization_benchmark(blk: Block, phi0, params: ReinitParams, out=None, work=None):
    """reinitialization_benchmark.F90:1-228 (M_Grad = 0, no obstacles, non-hybrid)."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    wp, wn, keep = _ws(blk, _lib.LSM_OP_REINIT_GODUNOV, work)
    p = params.c()
    check(L.lsm_reinit_godunov(blk.ref, C.byref(p), _f64(phi0, blk.shape, "phi0"),
                               _f64(out, blk.shape, "out"), wp, wn, _stream()), "lsm_reinit_godunov")
    return out


# ---------------------------------------------------------------------------
# water volume and the full step
# ---------------------------------------------------------------------------

def TotalWaterVolume(blk: Block, phi, with_mask: bool = True, work=None):
    """TotalWaterVolume.F90:1-297: returns (volumes, AdvectionNodes) where
    volumes = [K-tetrahedra, second tetrahedra, bulk cells] (CUDA float64)."""
    L = _lib.load()
    vol = torch.zeros(3, dtype=torch.float64, device=blk.device)
    mask = blk.empty(torch.int32) if with_mask else None
    wp, wn, keep = _ws(blk, _lib.LSM_OP_WATER_VOLUME, work)
    check(L.lsm_water_volume(blk.ref, _f64(phi, blk.shape, "phi"), C.c_void_p(vol.data_ptr()),
                             _i32(mask, blk.shape, "mask"), wp, wn, _stream()), "lsm_water_volume")
    return vol, mask


def levelsetmethod(blk: Block, phi, phi_n, u, v, w, params: ReinitParams, delti_lsm: float,
                   numiter_lsm: int = 1, call_levelsetmethod: bool = True,
                   call_reinitialisation: bool = True, RNfreq: int = 1, ntime: int = 1,
                   mode="fast", want_volume: bool = True, work=None):
    """levelsetmethod.F90:1-372: phi is overwritten with phi^(n+1); returns
    the pre-advection water volumes (or None).  Sussman reinitialisation, or
    with params.hybrid_reinitialisation the geometric one, the advection then
    restricted to the AdvectionNodes band."""
    L = _lib.load()
    vol = torch.zeros(3, dtype=torch.float64, device=blk.device) if want_volume else None
    wp, wn, keep = _ws(blk, _lib.LSM_OP_LEVELSET_STEP, work)
    p = params.c()
    check(L.lsm_levelset_step(blk.ref, C.byref(p), _f64(phi, blk.shape, "phi"),
                              _f64(phi_n, blk.shape, "phi_n"), _f64(u, blk.shape, "u"),
                              _f64(v, blk.shape, "v"), _f64(w, blk.shape, "w"), 1, int(numiter_lsm),
                              float(delti_lsm), int(bool(call_levelsetmethod)),
                              int(bool(call_reinitialisation)), int(RNfreq), int(ntime), MODES[mode],
                              C.c_void_p(vol.data_ptr()) if vol is not None else None, wp, wn,
                              _stream()), "lsm_levelset_step")
    return vol


def levelset_step_adjoint(blk: Block, phi_n, ucn, delti_lsm: float, params: ReinitParams, adj_np1,
                          mode="fast", adj_ucn=None, out=None, phi_np1=None, work=None):
    """Reverse mode of one full step phi^(n+1) = bcond_lsm(reinitialisation_test2(
    ad_RKTVD3(phi^n))) (levelsetmethod with numiter_lsm = 1, reinitialisation
    every step): returns dJ/dphi^n and ADDS dJ/d(U/J) to adj_ucn when given;
    phi_np1 (optional) receives the forward result."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    if work is None:
        work = blk.adjoint_workspace(_lib.LSM_OP_STEP_ADJOINT, params.RNitermax)
    au = [_f64(t, blk.shape, "adj_ucn") for t in adj_ucn] if adj_ucn is not None else [None] * 3
    p = params.c()
    check(L.lsm_levelset_step_adjoint(blk.ref, C.byref(p), _f64(phi_n, blk.shape, "phi_n"),
                                      *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm),
                                      MODES[mode], _f64(adj_np1, blk.shape, "adj_np1"),
                                      _f64(out, blk.shape, "out"), *au,
                                      _f64(phi_np1, blk.shape, "phi_np1") if phi_np1 is not None else None,
                                      C.c_void_p(work.data_ptr()), work.numel(), _stream()),
          "lsm_levelset_step_adjoint")
    return out


def levelset_step_record(blk: Block, phi_n, ucn, delti_lsm: float, params: ReinitParams, mode="fast",
                         phi_np1=None, trace=None):
    """Forward step recorded for its adjoint: returns (phi^(n+1), trace)."""
    L = _lib.load()
    if phi_np1 is None:
        phi_np1 = blk.empty()
    if trace is None:
        trace = blk.adjoint_workspace(_lib.LSM_OP_STEP_TRACE, params.RNitermax)
    p = params.c()
    check(L.lsm_levelset_step_record(blk.ref, C.byref(p), _f64(phi_n, blk.shape, "phi_n"),
                                     *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                     _f64(phi_np1, blk.shape, "phi_np1"), C.c_void_p(trace.data_ptr()),
                                     trace.numel(), _stream()), "lsm_levelset_step_record")
    return phi_np1, trace


def levelset_step_adjoint_recorded(blk: Block, phi_n, ucn, delti_lsm: float, params: ReinitParams, trace,
                                   adj_np1, mode="fast", adj_ucn=None, out=None, work=None):
    """Reverse sweep of a recorded step: dJ/dphi^n (dJ/d(U/J) added to adj_ucn)."""
    L = _lib.load()
    if out is None:
        out = blk.empty()
    if work is None:
        work = blk.adjoint_workspace(_lib.LSM_OP_STEP_ADJOINT_RECORDED, params.RNitermax)
    au = [_f64(t, blk.shape, "adj_ucn") for t in adj_ucn] if adj_ucn is not None else [None] * 3
    p = params.c()
    check(L.lsm_levelset_step_adjoint_recorded(blk.ref, C.byref(p), _f64(phi_n, blk.shape, "phi_n"),
                                               *[_f64(t, blk.shape, "ucn") for t in ucn], float(delti_lsm),
                                               MODES[mode], C.c_void_p(trace.data_ptr()), trace.numel(),
                                               _f64(adj_np1, blk.shape, "adj_np1"), _f64(out, blk.shape, "out"),
                                               *au, C.c_void_p(work.data_ptr()), work.numel(), _stream()),
          "lsm_levelset_step_adjoint_recorded")
    return out
