"""Level-set parameters from the reference's input files.

The reference configures the level-set method in two steps: the launcher
reads `control.dat` (`key = value` lines, simulation_setup_run.py:88-114) and
writes `ind3dmg_LSM.dat` (simulation_setup_run.py:239-329), which the solver
reads with list-directed Fortran reads (init_LSM.F90:129-183) into the
variables of module global_lsm.  This module restates both steps so a case
set up for the reference runs here with the same parameters:

    p = read_control("control.dat")          # dict, launcher semantics
    text = ind3dmg_lsm_text(p)               # the file the launcher writes
    prm = read_ind3dmg_lsm(text, delti=float(p["dt"]))
    prm.check_supported()
    reinit = prm.reinit_params()             # lsm_amd.api.ReinitParams

Field names follow global_lsm.F90.  No torch import happens here.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

# value maps of write_ind3dmg_LSM (simulation_setup_run.py:258-282)
BOUNDARY_TYPES = {"absorption": 0, "extrapolation": 1, "advection": 2}
EXTRAPOLATION_OPTIONS = {"upwind": 1, "central_difference": 0}
NARROW_BAND_OPTIONS = {"off": 0, "on": 1}
ADVECTION_SCHEMES = {"UPWIND": 0, "WENO3": 1}
SUSSMAN_CORRECTION_METHODS = {"point_constant": 0, "simpson_rule": 1}


def read_control(path: str) -> Dict[str, str]:
    """simulation_setup_run.py:88-114: every line holding '=' gives
    key = text left of the first '=' and value = text between the first and
    second '=', both stripped (comment lines with '=' included, as there)."""
    params: Dict[str, str] = {}
    with open(path) as f:
        for line in f:
            parts = line.split("=")
            if len(parts) > 1:
                params[parts[0].strip()] = parts[1].strip()
    return params


def _debug_steps(p: Dict[str, str]) -> List[int]:
    # simulation_setup_run.py:243-254
    s = p["LS_time_steps_to_save"]
    if s == "0":
        return []
    return [int(t.strip(" ")) for t in s.split(",")]


def ind3dmg_lsm_text(p: Dict[str, str]) -> str:
    """The ind3dmg_LSM.dat text write_ind3dmg_LSM (simulation_setup_run.py:
    286-324) produces from the control.dat parameters, line for line."""
    steps = _debug_steps(p)
    bc = " ".join(str(BOUNDARY_TYPES[p[k]]) for k in ("i1_LSM", "im_LSM", "j1_LSM", "jm_LSM",
                                                        "k1_LSM", "km_LSM"))
    lines = [
        f"{p['epsilon']} {p['level_set_advection_iterations']}",
        "0.0",
        f"{p['rho_air/rho_ref']} {p['mu_air/mu_ref']} {p['pressure_dissipation_LSM_air']} "
        f"{p['pressure_dissipation_LSM_ref']}",
        f"{p['Froude_number']} {p['Weber_number']}",
        "0 700 0.005 0",
        f"{p['reinitialization_iterations']} {p['reinitialization_dtau']} "
        f"{p['reinitilization_iteration_frequency']}",
        f"{p['phi_save_iterations']}",
        "0",
        bc,
        f"{EXTRAPOLATION_OPTIONS[p['extrapolation_option']]}",
        f"{NARROW_BAND_OPTIONS[p['narrow_band']]} {p['narrow_band_thickness']}",
        f"{ADVECTION_SCHEMES[p['advection_scheme']]}",
        f"{p['number_of_obstacles']}",
        f"{SUSSMAN_CORRECTION_METHODS[p['sussman_correction_method']]}",
        f"{p['call_levelsetmethod']}",
        f"{p['call_reinitialisation']}",
        f"{p['hybrid_reinitialisation']}",
        f"{p['sweep_lsqm']}",
        f"{p['radius_lsqm']}",
        f"{p['ConvergenceToleranceGeomReini']}",
        f"{p['TotalVolumeComputation']}",
        f"{p['GlobalMassCorrection']}",
        f"{len(steps)}",
    ]
    lines += [f"{t}" for t in steps]
    lines += [
        f"{p['epsReinitialisation']}",
        f"{p['OrderLSAdvectionBoundaries']}",
        f"{p['OrderReinitialisationBoundaries']}",
        f"{p['ENOBCReinitialisation']}",
        f"{p['BigPhi']}",
        f"{p['limit_ghost_velocities']}",
        f"{p['zero_pressure_fs']}",
    ]
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------------------
# list-directed input (the subset the file uses): values separated by blanks
# or commas, as many records as a READ needs, the rest of its last record
# skipped; r*c repeat counts; '/' ends the record; logicals by their first
# letter after an optional '.'; D exponents in reals
# ---------------------------------------------------------------------------
class _ListDirected:
    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def read(self, n: int) -> List[str]:
        vals: List[str] = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise EOFError(f"end of file while reading {n} values (got {len(vals)})")
            line = self.lines[self.pos]
            self.pos += 1
            slash = line.find("/")
            stop = slash >= 0
            if stop:
                line = line[:slash]
            for tok in re.split(r"[,\s]+", line.strip()):
                if not tok:
                    continue
                m = re.fullmatch(r"(\d+)\*(.*)", tok)
                if m:
                    vals.extend([m.group(2)] * int(m.group(1)))
                else:
                    vals.append(tok)
            if stop:
                break
        if len(vals) < n:
            # a '/' would leave the remaining variables unchanged; the level-set
            # file never relies on that, so it is reported instead
            raise ValueError(f"'/' ended a read after {len(vals)} of {n} values")
        return vals[:n]


def _int(tok: str) -> int:
    return int(tok)


def _real(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def _logical(tok: str) -> bool:
    t = tok.lstrip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


@dataclass
class LsmParams:
    """global_lsm.F90 variables set by init_input_fileLSM (init_LSM.F90:136-181)."""
    epslsm: float = 0.0
    numiter_lsm: int = 1
    thetainc: float = 0.0
    drho: float = 0.0
    dmu: float = 0.0
    diss_w: float = 0.0
    diss_a: float = 0.0
    FrLSM: float = 0.0
    WeLSM: float = 0.0
    IG: int = 0
    IGitermax: float = 0.0
    IGtimestep: float = 0.0
    IGstep2: int = 0
    RNitermax: int = 0
    RNtimestep: float = 0.0
    RNfreq: int = 1
    phi_outputiter: int = 0
    k_surface: int = 0
    btype_lsm: List[tuple] = field(default_factory=list)     # per zone: (i1, im, j1, jm, k1, km)
    bc_extrapolation_order: int = 1
    isnarrow: int = 0
    narrowcoef: float = 0.0
    orderAD: int = 1
    nobstacles: int = 0
    sussman_correction_method: int = 0
    call_levelsetmethod: bool = True
    call_reinitialisation: bool = True
    hybrid_reinitialisation: bool = False
    sweep_lsqm: int = 0
    radius_lsqm: float = 0.0
    ConvergenceToleranceGeomReini: float = 0.0
    TotalVolumeComputation: bool = False
    GlobalMassCorrection: bool = False
    save_LSdebug_tsteps: List[int] = field(default_factory=list)
    epsReinitialisation: float = 0.0
    OrderLSAdvectionBoundaries: int = 1
    OrderReinitialisationBoundaries: int = 1
    ENOBCReinitialisation: bool = False
    BigPhi: float = 5.0
    limit_ghost_velocities: bool = False
    zero_pressure_fs: bool = False
    delti_lsm: Optional[float] = None

    def check_supported(self, zone: int = 1) -> None:
        """Raises NotImplementedError for options outside the GPU path (the
        C ABI rejects them the same way)."""
        bad = []
        if self.orderAD not in (0, 1):
            bad.append(f"orderAD = {self.orderAD} (0: first-order upwind, 1: WENO3)")
        if self.sussman_correction_method not in (0, 1):
            bad.append(f"sussman_correction_method = {self.sussman_correction_method} (0: point, 1: filtered)")
        if self.hybrid_reinitialisation and self.call_reinitialisation and self.GlobalMassCorrection:
            bad.append("hybrid reinitialisation with GlobalMassCorrection (global mass correction)")
        if self.nobstacles != 0:
            bad.append(f"{self.nobstacles} obstacles")
        if self.OrderReinitialisationBoundaries not in (1, 2, 3):
            bad.append(f"OrderReinitialisationBoundaries = {self.OrderReinitialisationBoundaries}")
        if self.btype_lsm and any(t not in (0, 1, 2) for t in self.btype_lsm[zone - 1]):
            bad.append(f"btype_lsm = {self.btype_lsm[zone - 1]}")
        if bad:
            raise NotImplementedError("unsupported level-set configuration: " + "; ".join(bad))

    def reinit_params(self, M_Sphi0: int = 1, M_TI: int = 1):
        """The reinitialisation parameters as lsm_amd.api.ReinitParams."""
        from . import api
        return api.ReinitParams(epslsm=self.epslsm, RNtimestep=self.RNtimestep,
                                RNitermax=self.RNitermax,
                                epsReinitialisation=self.epsReinitialisation,
                                OrderReinitialisationBoundaries=self.OrderReinitialisationBoundaries,
                                ENOBCReinitialisation=self.ENOBCReinitialisation,
                                M_Sphi0=M_Sphi0, M_TI=M_TI,
                                hybrid_reinitialisation=self.hybrid_reinitialisation,
                                sussman_correction_method=self.sussman_correction_method)


def read_ind3dmg_lsm(src: str, nzone: int = 1, delti: Optional[float] = None) -> LsmParams:
    """init_input_fileLSM (init_LSM.F90:129-183): the reads of lines 136-168
    in order, then delti_lsm = delti / numiter_lsm (line 181) when the
    physical time step `delti` is given.  `src` is a path or the file text."""
    text = open(src).read() if os.path.exists(src) else src
    r = _ListDirected(text)
    p = LsmParams()
    v = r.read(2); p.epslsm, p.numiter_lsm = _real(v[0]), _int(v[1])
    p.thetainc = _real(r.read(1)[0])
    v = r.read(4); p.drho, p.dmu, p.diss_w, p.diss_a = map(_real, v)
    v = r.read(2); p.FrLSM, p.WeLSM = map(_real, v)
    v = r.read(4); p.IG, p.IGitermax, p.IGtimestep, p.IGstep2 = _int(v[0]), _real(v[1]), _real(v[2]), _int(v[3])
    v = r.read(3); p.RNitermax, p.RNtimestep, p.RNfreq = _int(v[0]), _real(v[1]), _int(v[2])
    p.phi_outputiter = _int(r.read(1)[0])
    p.k_surface = _int(r.read(1)[0])
    v = [_int(t) for t in r.read(6 * nzone)]
    p.btype_lsm = [tuple(v[6 * z:6 * z + 6]) for z in range(nzone)]
    p.bc_extrapolation_order = _int(r.read(1)[0])
    v = r.read(2); p.isnarrow, p.narrowcoef = _int(v[0]), _real(v[1])
    p.orderAD = _int(r.read(1)[0])
    p.nobstacles = _int(r.read(1)[0])
    p.sussman_correction_method = _int(r.read(1)[0])
    p.call_levelsetmethod = _logical(r.read(1)[0])
    p.call_reinitialisation = _logical(r.read(1)[0])
    p.hybrid_reinitialisation = _logical(r.read(1)[0])
    p.sweep_lsqm = _int(r.read(1)[0])
    p.radius_lsqm = _real(r.read(1)[0])
    p.ConvergenceToleranceGeomReini = _real(r.read(1)[0])
    p.TotalVolumeComputation = _logical(r.read(1)[0])
    p.GlobalMassCorrection = _logical(r.read(1)[0])
    n = _int(r.read(1)[0])
    if n > 20:
        raise ValueError("at most 20 debug time steps (save_LSdebug_tsteps has 20 entries)")
    p.save_LSdebug_tsteps = [_int(t) for t in r.read(n)] if n > 0 else []
    p.epsReinitialisation = _real(r.read(1)[0])
    p.OrderLSAdvectionBoundaries = _int(r.read(1)[0])
    p.OrderReinitialisationBoundaries = _int(r.read(1)[0])
    p.ENOBCReinitialisation = _logical(r.read(1)[0])
    p.BigPhi = _real(r.read(1)[0])
    p.limit_ghost_velocities = _logical(r.read(1)[0])
    p.zero_pressure_fs = _logical(r.read(1)[0])
    if delti is not None:
        p.delti_lsm = float(delti) / float(p.numiter_lsm)
    return p


def from_control(path: str, nzone: int = 1) -> LsmParams:
    """control.dat -> ind3dmg_LSM.dat text -> parameters, with the physical
    time step `dt` of control.dat giving delti_lsm."""
    p = read_control(path)
    return read_ind3dmg_lsm(ind3dmg_lsm_text(p), nzone=nzone, delti=_real(p["dt"]))


def btype_for_block(prm: LsmParams, zone: int = 1) -> Sequence[int]:
    """btype_lsm(1:6, zone) in the order of lsm_amd.api.Block's btype."""
    return tuple(prm.btype_lsm[zone - 1])
