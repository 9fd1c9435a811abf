"""CPU tests of the control.dat -> ind3dmg_LSM.dat -> parameter mapping
(lsm_amd.control), restating simulation_setup_run.py:88-114, 239-329 and
init_LSM.F90:129-183."""
import os

import pytest

from lsm_amd import control as C

CONTROL = """\
!----------------------------------------------------------
! LEVEL SET (comment lines with = are parsed too: a = b)
!----------------------------------------------------------
dt = 0.002
epsilon = 0.015
rho_air/rho_ref = 0.0012
mu_air/mu_ref = 0.018
pressure_dissipation_LSM_air = 0.1
pressure_dissipation_LSM_ref = 0.2
Weber_number = 0.0033
Froude_number = 0.75
BigPhi = 4.5
i1_LSM = extrapolation
im_LSM = advection
j1_LSM = absorption
jm_LSM = extrapolation
k1_LSM = extrapolation
km_LSM = absorption
extrapolation_option = upwind
level_set_advection_iterations = 4
reinitialization_iterations = 12
reinitialization_dtau = 0.0005
epsReinitialisation = 1.0E-05
reinitilization_iteration_frequency = 2
phi_save_iterations = 100
zero_pressure_fs = .false.
limit_ghost_velocities = .true.
LS_time_steps_to_save = 5, 7,11
narrow_band = on
narrow_band_thickness = 3.0
advection_scheme = WENO3
OrderLSAdvectionBoundaries = 2
OrderReinitialisationBoundaries = 2
ENOBCReinitialisation = .true.
sussman_correction_method = point_constant
number_of_obstacles = 0
call_levelsetmethod = .true.
call_reinitialisation = .true.
hybrid_reinitialisation = .false.
ConvergenceToleranceGeomReini = 1.0E-09
sweep_lsqm = 3
radius_lsqm = 3.5
    TotalVolumeComputation = .false.
    GlobalMassCorrection = .true.
"""


@pytest.fixture
def ctl(tmp_path):
    p = tmp_path / "control.dat"
    p.write_text(CONTROL)
    return str(p)


def test_read_control_semantics(ctl, tmp_path):
    p = C.read_control(ctl)
    assert p["dt"] == "0.002"
    assert p["rho_air/rho_ref"] == "0.0012"
    assert p["TotalVolumeComputation"] == ".false."          # leading blanks stripped
    assert p["LS_time_steps_to_save"] == "5, 7,11"
    # a comment line holding '=' becomes an entry whose value is the text between
    # the first and second '='
    assert p["! LEVEL SET (comment lines with"] == "are parsed too: a"
    f = tmp_path / "two.dat"
    f.write_text("x = 1 = 2\n")
    assert C.read_control(str(f))["x"] == "1"


def test_ind3dmg_lsm_text_layout(ctl):
    lines = C.ind3dmg_lsm_text(C.read_control(ctl)).splitlines()
    # 23 fixed records before the debug list, 3 debug steps, 7 trailing records
    assert len(lines) == 23 + 3 + 7
    assert lines[0] == "0.015 4"
    assert lines[4] == "0 700 0.005 0"
    assert lines[5] == "12 0.0005 2"
    assert lines[8] == "1 2 0 1 1 0"                           # absorption 0, extrapolation 1, advection 2
    assert lines[10] == "1 3.0"
    assert lines[11] == "1"                                    # WENO3
    assert lines[22] == "3" and lines[23:26] == ["5", "7", "11"]
    assert lines[-1] == ".false." and lines[-3] == "4.5"


def test_from_control(ctl):
    prm = C.from_control(ctl)
    assert prm.epslsm == 0.015 and prm.numiter_lsm == 4
    assert prm.RNitermax == 12 and prm.RNtimestep == 0.0005 and prm.RNfreq == 2
    assert prm.btype_lsm == [(1, 2, 0, 1, 1, 0)]
    assert prm.orderAD == 1 and prm.sussman_correction_method == 0
    assert prm.call_levelsetmethod and prm.call_reinitialisation and not prm.hybrid_reinitialisation
    assert prm.ENOBCReinitialisation and prm.OrderReinitialisationBoundaries == 2
    assert prm.GlobalMassCorrection and not prm.TotalVolumeComputation
    assert prm.save_LSdebug_tsteps == [5, 7, 11]
    assert prm.BigPhi == 4.5 and prm.limit_ghost_velocities and not prm.zero_pressure_fs
    assert prm.delti_lsm == 0.002 / 4                          # init_LSM.F90:181
    prm.check_supported()
    assert C.btype_for_block(prm) == (1, 2, 0, 1, 1, 0)


def test_list_directed_forms():
    """Values split over records or packed on one, commas, r*c repeats, D
    exponents, T/F logicals, trailing text ignored, '/' ending a record."""
    text = "\n".join([
        "1.5d-2, 3   trailing words are skipped",
        "0.0",
        "1 2",
        "3 4",
        "0.5 0.25",
        "0 700 0.005 0",
        "10 1.0D-3 1",
        "5", "0",
        "6*1",
        "1", "0 2.5", "1", "0", "0",
        "T", ".TRUE.", "f",
        "3", "3.5", "1e-9", "F", ".false.",
        "0",
        "1.0E-05", "2", "1", "F", "5.0", "F", "F / the rest of the record is skipped",
    ])
    prm = C.read_ind3dmg_lsm(text, delti=0.03)
    assert prm.epslsm == 1.5e-2 and prm.numiter_lsm == 3
    assert (prm.drho, prm.dmu, prm.diss_w, prm.diss_a) == (1.0, 2.0, 3.0, 4.0)
    assert prm.RNtimestep == 1.0e-3
    assert prm.btype_lsm == [(1, 1, 1, 1, 1, 1)]
    assert prm.call_levelsetmethod and prm.call_reinitialisation and not prm.hybrid_reinitialisation
    assert prm.save_LSdebug_tsteps == []
    assert prm.zero_pressure_fs is False and prm.delti_lsm == 0.01
    with pytest.raises(ValueError):
        C.read_ind3dmg_lsm(text.replace("10 1.0D-3 1", "10 /"))


def test_unsupported_options_are_rejected(ctl):
    prm = C.from_control(ctl)
    prm.hybrid_reinitialisation = True          # geometric reinitialisation: supported
    prm.GlobalMassCorrection = False
    prm.check_supported()
    assert prm.reinit_params().hybrid_reinitialisation
    prm.GlobalMassCorrection = True             # ... without the global mass correction
    with pytest.raises(NotImplementedError, match="GlobalMassCorrection"):
        prm.check_supported()
    prm.hybrid_reinitialisation = False
    prm.GlobalMassCorrection = False
    prm.orderAD = 0                    # the first-order upwind right-hand side is supported
    prm.check_supported()
    prm.orderAD = 2
    with pytest.raises(NotImplementedError, match="orderAD"):
        prm.check_supported()
    prm.orderAD = 1
    prm.sussman_correction_method = 1  # the filtered lambda (Godunov path) is supported
    prm.check_supported()
    assert prm.reinit_params().c().lambda_method == 1
    prm.sussman_correction_method = 2
    with pytest.raises(NotImplementedError, match="sussman_correction_method"):
        prm.check_supported()


REF_CONTROL = "/root/reference/control.dat"


@pytest.mark.skipif(not os.path.exists(REF_CONTROL), reason="reference tree not present")
def test_reference_control_file_maps():
    """The reference's own control.dat goes through the same mapping; its
    configuration (hybrid geometric reinitialisation without the global mass
    correction) is supported."""
    prm = C.from_control(REF_CONTROL)
    assert prm.numiter_lsm >= 1 and prm.delti_lsm > 0
    assert len(prm.btype_lsm[0]) == 6
    assert prm.hybrid_reinitialisation and not prm.GlobalMassCorrection
    prm.check_supported()
    assert prm.reinit_params().hybrid_reinitialisation
