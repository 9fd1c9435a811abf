"""ctypes binding of liblsm_hip.so (the C ABI declared in include/lsm_hip.h).

The library must be built (python -m lsm_amd.build); loading fails loudly
otherwise -- there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblsm_hip.so")

LSM_OK = 0
LSM_ERR_ARG, LSM_ERR_UNSUPPORTED, LSM_ERR_HIP, LSM_ERR_WORKSPACE = -1, -2, -3, -4
LSM_FACE_ALL = 0x3F
LSM_ADVECT_FAST = 0
LSM_ADVECT_EXACT = 1
LSM_ADVECT_UPWIND1 = 2
LSM_OP_RK3, LSM_OP_REINIT_ENO2, LSM_OP_REINIT_GODUNOV, LSM_OP_WATER_VOLUME, LSM_OP_LEVELSET_STEP = 1, 2, 3, 4, 5
LSM_OP_RK3_ADJOINT = 6
LSM_OP_REINIT_ENO2_ADJOINT, LSM_OP_STEP_ADJOINT, LSM_OP_STEP_TRACE, LSM_OP_STEP_ADJOINT_RECORDED = 7, 8, 9, 10
LSM_OP_REINIT_GEOMETRIC = 11
LSM_OP_LAMBDA_RN_SIMPSON = 12

_vp = C.c_void_p


LSM_MAX_BLANK = 4


class LsmBlank(C.Structure):
    _fields_ = [("ia", C.c_int32), ("ib", C.c_int32), ("ja", C.c_int32), ("jb", C.c_int32),
                ("ka", C.c_int32), ("kb", C.c_int32), ("blktype", C.c_int32 * 4)]


class LsmBlock(C.Structure):
    _fields_ = [("il", C.c_int32), ("iu", C.c_int32), ("jl", C.c_int32), ("ju", C.c_int32),
                ("kl", C.c_int32), ("ku", C.c_int32),
                ("igp", C.c_int32), ("jgp", C.c_int32), ("kgp", C.c_int32),
                ("phys_faces", C.c_int32), ("btype", C.c_int32 * 6),
                ("dc", C.c_double), ("de", C.c_double), ("dz", C.c_double),
                ("big_phi", C.c_double),
                ("csi", _vp * 3), ("eta", _vp * 3), ("zet", _vp * 3),
                ("aj", _vp), ("x", _vp), ("y", _vp), ("z", _vp),
                ("nblk", C.c_int32), ("blank", LsmBlank * LSM_MAX_BLANK), ("gten", _vp)]


class LsmReinitParams(C.Structure):
    _fields_ = [("epslsm", C.c_double), ("eps_reinit", C.c_double), ("rn_timestep", C.c_double),
                ("rn_itermax", C.c_int32), ("order_bc", C.c_int32), ("eno_bc", C.c_int32),
                ("m_sphi0", C.c_int32), ("m_ti", C.c_int32), ("hybrid", C.c_int32),
                ("lambda_method", C.c_int32)]


class LsmBatchBlock(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32), ("lattice", C.c_int32 * 3),
                ("phys_faces", C.c_int32), ("btype", C.c_int32 * 6), ("big_phi", C.c_double),
                ("dc", C.c_double), ("de", C.c_double), ("dz", C.c_double), ("offset", C.c_int64),
                ("pitch", C.c_int32)]


class LsmBatchDesc(C.Structure):
    _fields_ = [("nblocks", C.c_int32), ("ghost", C.c_int32), ("connected", C.c_int32),
                ("lattice", C.c_int32 * 3), ("blocks", C.POINTER(LsmBatchBlock)), ("aj", _vp),
                ("csi", _vp * 3), ("eta", _vp * 3), ("zet", _vp * 3), ("total", C.c_int64)]


LSM_BATCH_QUEUE, LSM_BATCH_GRID, LSM_BATCH_PERBLOCK = 0, 1, 2
_pp = C.POINTER(_vp)

# name -> (restype, argtypes)
_SIGS = {
    "lsm_batch_create": (C.c_int, [C.POINTER(LsmBatchDesc), C.POINTER(_vp)]),
    "lsm_batch_destroy": (C.c_int, [_vp]),
    "lsm_batch_set_config": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "lsm_batch_get_config": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "lsm_batch_set_reinit_tile": (C.c_int, [_vp, C.c_int]),
    "lsm_batch_get_reinit_tile": (C.c_int, [_vp]),
    "lsm_batch_workspace_bytes": (C.c_size_t, [_vp, C.c_int]),
    "lsm_batch_halo_fill": (C.c_int, [_vp, _vp, _vp]),
    "lsm_batch_contra_velocity": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lsm_batch_reinit_sussman_eno2": (C.c_int, [_vp, C.POINTER(LsmReinitParams), _vp, _vp, _vp, _vp, C.c_size_t,
                                                _vp]),
    "lsm_batch_bcond": (C.c_int, [_vp, _vp, _vp]),
    "lsm_batch_levelset_step": (C.c_int, [_vp, C.POINTER(LsmReinitParams), _vp, _vp, _vp, _vp, _vp, C.c_double,
                                          C.c_int, _vp, C.c_size_t, _vp]),
    "lsm_batch_rk3_step": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp, C.c_size_t, _vp]),
    "lsm_batch_autotune": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp, C.c_size_t, C.c_int,
                                     C.c_int, C.POINTER(C.c_double), _vp]),
    "lsm_batch_graph_capture": (C.c_int, [_vp, C.POINTER(LsmReinitParams), _vp, _vp, _vp, _vp, _vp, C.c_double,
                                          C.c_int, _vp, C.c_size_t]),
    "lsm_batch_graph_launch": (C.c_int, [_vp, _vp]),
    "lsm_version": (C.c_char_p, []),
    "lsm_last_error": (C.c_char_p, []),
    "lsm_workspace_bytes": (C.c_size_t, [C.POINTER(LsmBlock), C.c_int]),
    "lsm_contra_velocity": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp]),
    "lsm_weno3_rhs": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "lsm_upwind_rhs": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lsm_bcond": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp]),
    "lsm_rk3_step": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int,
                               _vp, C.c_size_t, _vp]),
    "lsm_advect_fold": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    "lsm_rk3_step_folded": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t,
                                      _vp]),
    "lsm_rk3_step_adjoint": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int,
                                       _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "lsm_adjoint_workspace_bytes": (C.c_size_t, [C.POINTER(LsmBlock), C.c_int, C.c_int]),
    "lsm_bcond_adjoint": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp]),
    "lsm_reinit_sussman_eno2_adjoint": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp,
                                                  _vp, _vp, _vp, C.c_size_t, _vp]),
    "lsm_levelset_step_adjoint": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp, _vp,
                                            _vp, C.c_double, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            C.c_size_t, _vp]),
    "lsm_levelset_step_record": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp, _vp, _vp,
                                           C.c_double, C.c_int, _vp, _vp, C.c_size_t, _vp]),
    "lsm_levelset_step_adjoint_recorded": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp, _vp,
                                                     _vp, C.c_double, C.c_int, _vp, C.c_size_t, _vp, _vp, _vp, _vp,
                                                     _vp, _vp, C.c_size_t, _vp]),
    "lsm_normgrad_eno2": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, C.c_int, C.c_int, _vp, _vp]),
    "lsm_phi_gradient_eno2": (C.c_int, [C.POINTER(LsmBlock), _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "lsm_h_gradient_eno2": (C.c_int, [C.POINTER(LsmBlock), _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
    "lsm_calc_h": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp]),
    "lsm_free_surface": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp]),
    "lsm_free_surface_adjoint": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, _vp]),
    "lsm_metric_tensor": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp]),
    "lsm_reinit_sussman_eno2": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp,
                                          _vp, _vp, C.c_size_t, _vp]),
    "lsm_sign_ini": (C.c_int, [C.POINTER(LsmBlock), _vp, C.c_double, C.c_int, _vp, _vp, _vp]),
    "lsm_rh_rn": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp]),
    "lsm_lambda_rn": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp]),
    "lsm_lambda_rn_simpson": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp,
                                        C.c_size_t, _vp]),
    "lsm_reinit_geometric": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "lsm_reinit_godunov": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp, _vp,
                                     C.c_size_t, _vp]),
    "lsm_water_volume": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "lsm_water_volume_noalloc": (C.c_int, [C.POINTER(LsmBlock), _vp, _vp, _vp, _vp]),
    "lsm_levelset_step": (C.c_int, [C.POINTER(LsmBlock), C.POINTER(LsmReinitParams), _vp, _vp, _vp, _vp,
                                    _vp, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, _vp, _vp, C.c_size_t, _vp]),
    "lsm_timing_enable": (C.c_int, [C.c_int]),
    "lsm_timing_get": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    # host-memory adapter (arrays in host memory, synchronous)
    "lsm_host_setup": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.c_double,
                                 C.c_double, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lsm_host_release": (C.c_int, []),
    "lsm_host_invalidate": (C.c_int, []),
    "lsm_host_blanking": (C.c_int, [C.c_int32, _vp]),
    "lsm_multi_begin": (C.c_int, []),
    "lsm_multi_instance": (C.c_int, []),
    "lsm_multi_end": (C.c_int, [_vp]),
    "lsm_multi_cancel": (C.c_int, []),
    "lsm_host_consumer_blanking": (C.c_int, [C.c_int32, _vp]),
    "lsm_host_levelset_step": (C.c_int, [_vp, _vp, _vp, C.POINTER(LsmReinitParams), C.c_int, C.c_double,
                                         C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "lsm_host_contra_velocity": (C.c_int, [_vp, _vp]),
    "lsm_host_weno3_rhs": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "lsm_host_upwind_rhs": (C.c_int, [_vp, _vp, _vp]),
    "lsm_host_rk3": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int]),
    "lsm_host_bcond": (C.c_int, [_vp]),
    "lsm_host_normgrad": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp]),
    "lsm_host_reinit_eno2": (C.c_int, [_vp, _vp, C.POINTER(LsmReinitParams), C.POINTER(C.c_int32)]),
    "lsm_host_sign_ini": (C.c_int, [_vp, C.c_double, C.c_int, _vp, _vp]),
    "lsm_host_rh_rn": (C.c_int, [_vp, _vp, _vp]),
    "lsm_host_lambda_rn": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp]),
    "lsm_host_lambda_rn_simpson": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp]),
    "lsm_host_reinit_godunov": (C.c_int, [_vp, _vp, C.POINTER(LsmReinitParams)]),
    "lsm_host_water_volume": (C.c_int, [_vp, _vp, _vp]),
    "lsm_host_phi_gradient_eno2": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp,
                                             _vp, C.c_int, C.c_int, _vp]),
    "lsm_host_h_gradient_eno2": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp,
                                           _vp, C.c_int, C.c_int, _vp]),
    "lsm_host_calc_h": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp]),
    "lsm_host_reinit_geometric": (C.c_int, [_vp, _vp, _vp]),
}

EXPORTS = tuple(_SIGS)
LSM_KT_ADVECT_STAGE, LSM_KT_ADVECT_SHELL, LSM_KT_REINIT_STAGE, LSM_KT_VOLUME = 0, 1, 2, 3
LSM_KT_ADJOINT_STAGE = 4
LSM_KT_BATCH_STAGE = 5
LSM_KT_BATCH_REINIT = 6
LSM_KT_REINIT_ADJOINT = 7
LSM_KT_GEOMETRIC = 8

_lib = None


class LsmError(RuntimeError):
    pass


def load():
    """Loads liblsm_hip.so (importing torch first so the process uses one HIP
    runtime) and declares every exported signature."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401  (brings in the HIP runtime torch uses)
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise LsmError(f"{LIB_PATH} is not built: run `python -m lsm_amd.build` in "
                           "ns-lsm-single-phase_amd/ (there is no host fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str):
    if rc != LSM_OK:
        msg = load().lsm_last_error().decode(errors="replace")
        raise LsmError(f"{what} failed ({rc}): {msg}")
