"""CPU checks of the C ABI boundary (no GPU work is launched).

 - every function include/lsm_hip.h declares is exported by liblsm_hip.so and
   has a ctypes signature in lsm_amd._lib;
 - the ctypes structures and the Fortran bind(C) type have the C layout;
 - the product path fails loudly when the library is missing;
 - calls that fail argument validation return the documented error codes.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lsm_hip.h")
FORTRAN = os.path.join(ROOT, "ns-lsm-single-phase_amd", "fortran")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lsm_\w+)\s*\(", text, flags=re.M)))


def test_header_matches_bindings():
    from lsm_amd import _lib
    decl = declared_functions()
    assert len(decl) >= 30
    assert sorted(_lib.EXPORTS) == decl


def test_library_exports_every_declared_symbol():
    from lsm_amd import _lib
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert "gfx950" in lib.lsm_version().decode()


def test_struct_layout_matches_c(tmp_path):
    from lsm_amd._lib import LsmBlock, LsmReinitParams
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "lsm_hip.h"\n'
        "#define F(T, m) printf(#T \" \" #m \" %zu\\n\", offsetof(T, m))\n"
        "int main(void) {\n"
        '  printf("LsmBlock size %zu\\n", sizeof(LsmBlock));\n'
        '  printf("LsmReinitParams size %zu\\n", sizeof(LsmReinitParams));\n'
        + "".join(f"  F(LsmBlock, {f});\n" for f, _ in LsmBlock._fields_)
        + "".join(f"  F(LsmReinitParams, {f});\n" for f, _ in LsmReinitParams._fields_)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            t, m, v = line.split()
            got[(t, m)] = int(v)
    assert got[("LsmBlock", "size")] == C.sizeof(LsmBlock)
    assert got[("LsmReinitParams", "size")] == C.sizeof(LsmReinitParams)
    for T in (LsmBlock, LsmReinitParams):
        for f, _ in T._fields_:
            assert got[(T.__name__, f)] == getattr(T, f).offset, (T.__name__, f)


def test_fortran_interface_compiles_and_matches_c(tmp_path):
    """The Fortran bind(C) module compiles and its lsm_reinit_params type has
    the size of LsmReinitParams."""
    from lsm_amd import _lib
    from lsm_amd._lib import LsmReinitParams
    fc = shutil.which("amdflang")
    if fc is None:
        pytest.skip("amdflang not available")
    prog = tmp_path / "sz.F90"
    prog.write_text("program sz\n use iso_c_binding\n use lsm_hip_f\n type(lsm_reinit_params) :: p\n"
                    " print '(i0)', c_sizeof(p)\nend program sz\n")
    exe = tmp_path / "sz"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([fc, "-module-dir", str(tmp_path), "-I", str(tmp_path),
                    os.path.join(FORTRAN, "lsm_hip_f.F90"), str(prog), "-o", str(exe),
                    "-L" + libdir, "-Wl,-rpath," + libdir, "-llsm_hip"], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == C.sizeof(LsmReinitParams)


def test_shim_files_cover_the_hot_path():
    """One drop-in file per reference routine of the level-set path, each
    defining the routine of the same name."""
    routines = ["rhs_contra_j", "ad_RKTVD3", "calc_RH_AD_WENO3", "bcond_lsm", "calc_NormGrad_d",
                "reinitialisation_test2", "calc_sign_ini", "calc_RH_RN", "calc_lambda_RN",
                "reinitialization_benchmark", "TotalWaterVolume", "calc_phi_gradient_ENO2", "calc_h",
                "calc_h_gradient_ENO2", "calc_RH_AD"]
    for r in routines:
        text = open(os.path.join(FORTRAN, r + ".F90")).read()
        assert re.search(rf"^\s*subroutine\s+{r}\b", text, flags=re.M | re.I), r
        assert re.search(rf"end subroutine\s+{r}\b", text, flags=re.I), r


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from lsm_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "liblsm_hip.so"))
    with pytest.raises(_lib.LsmError, match="not built"):
        _lib.load()


def test_argument_errors_without_device():
    """The host adapter rejects calls before lsm_host_setup (no HIP call is
    made on that path) with LSM_ERR_ARG and a message."""
    from lsm_amd import _lib
    lib = _lib.load()
    rc = lib.lsm_host_bcond(None)
    assert rc == -1
    assert "lsm_host_setup" in lib.lsm_last_error().decode()
