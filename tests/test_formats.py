"""CPU tests of the reference file formats (lsm_amd.formats) and of the
metrics of a grid read from them."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fixtures import load
from lsm_amd import formats as F
from lsm_amd import inputs


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(4)]


def test_round_trip(tmp_path):
    s1, s2 = (5, 4, 6), (3, 2, 7)
    x1, y1, z1, d1 = _fields(s1, 1)
    x2, y2, z2, d2 = _fields(s2, 2)
    F.write_grid(str(tmp_path / "grid"), [(x1, y1, z1, d1), (x2, y2, z2, d2)])
    g = F.read_grid(str(tmp_path / "grid"), [s1, s2], turbulence=True)
    for a, b in zip(g[0] + g[1], (x1, y1, z1, d1, x2, y2, z2, d2)):
        assert np.array_equal(a, b)
    F.write_phi_ini(str(tmp_path / "phi_ini"), [(x1, y1), (x2, y2)])
    p = F.read_phi_ini(str(tmp_path / "phi_ini"), [s1, s2])
    assert np.array_equal(p[1][0], x2) and np.array_equal(p[1][1], y2)
    F.write_phi_restart(str(tmp_path / "phi.000"), z1, d1)
    a, b = F.read_phi_restart(str(tmp_path / "phi.000"), s1)
    assert np.array_equal(a, z1) and np.array_equal(b, d1)
    with pytest.raises(ValueError):
        F.read_phi_ini(str(tmp_path / "phi_ini"), [s2, s1])


def test_subrecords(tmp_path, monkeypatch):
    """Records longer than the subrecord limit are split and joined again."""
    monkeypatch.setattr(F, "_MAXREC", 64)
    x = np.arange(3 * 5 * 7, dtype=np.float64).reshape(3, 5, 7)
    F.write_phi_restart(str(tmp_path / "phi.001"), x, -x)
    raw = open(tmp_path / "phi.001", "rb").read()
    assert int.from_bytes(raw[:4], "little", signed=True) == -64   # continued
    a, b = F.read_phi_restart(str(tmp_path / "phi.001"), x.shape)
    assert np.array_equal(a, x) and np.array_equal(b, -x)


FORTRAN = """
program fmt
  implicit none
  integer, parameter :: ni = 5, nj = 4, nk = 3
  real(kind=8) :: x(ni,nj,nk), y(ni,nj,nk), z(ni,nj,nk), p(ni*nj*nk), q(ni*nj*nk)
  integer :: i, j, k, l
  do k = 1, nk
  do j = 1, nj
  do i = 1, ni
     x(i,j,k) = i + 10*j + 100*k
     y(i,j,k) = -x(i,j,k)
     z(i,j,k) = 0.5d0 * x(i,j,k)
  end do
  end do
  end do
  ! the reference's grid writer shape (init.F90:1872-1877, read there)
  open(41, file='grid', form='unformatted')
  write(41) (((x(i,j,k),i=1,ni),j=1,nj),k=1,nk)
  write(41) (((y(i,j,k),i=1,ni),j=1,nj),k=1,nk)
  write(41) (((z(i,j,k),i=1,ni),j=1,nj),k=1,nk)
  close(41)
  ! read a phi_ini written by Python (init_LSM.F90:350-357)
  open(42, file='phi_ini', form='unformatted')
  read(42) (p(l), l = 1, ni*nj*nk)
  read(42) (q(l), l = 1, ni*nj*nk)
  close(42)
  print '(2es25.16)', sum(p * [(real(l,8), l = 1, ni*nj*nk)]), sum(q)
end program fmt
"""


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_fortran_unformatted_compatibility(tmp_path):
    (tmp_path / "fmt.f90").write_text(FORTRAN)
    subprocess.run(["amdflang", "fmt.f90", "-o", "fmt"], cwd=tmp_path, check=True)
    shape = (3, 4, 5)
    rng = np.random.default_rng(5)
    phi, phi_n = rng.standard_normal(shape), rng.standard_normal(shape)
    F.write_phi_ini(str(tmp_path / "phi_ini"), [(phi, phi_n)])
    out = subprocess.run(["./fmt"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    s1, s2 = (float(t) for t in out.split())
    w = np.arange(1, phi.size + 1, dtype=np.float64)
    assert abs(s1 - np.sum(phi.ravel() * w)) <= 1e-12 * np.sum(np.abs(phi.ravel() * w))
    assert abs(s2 - np.sum(phi_n)) <= 1e-12 * np.sum(np.abs(phi_n))
    x, y, z = F.read_grid(str(tmp_path / "grid"), [shape])[0]
    kk, jj, ii = np.meshgrid(np.arange(1, 4), np.arange(1, 5), np.arange(1, 6), indexing="ij")
    ref = ii + 10.0 * jj + 100.0 * kk
    assert np.array_equal(x, ref) and np.array_equal(y, -ref) and np.array_equal(z, 0.5 * ref)


@pytest.mark.parametrize("name", ["curvilinear", "channel"])
def test_metrics_match_compiled_mg_metrics(name, tmp_path):
    """compute_metrics (used for grids read from file) is bit-identical to
    the compiled mg_metrics on the fixture grids; the grid goes through the
    file format on the way."""
    fx = load(name)
    g = fx.grid
    F.write_grid(str(tmp_path / "grid"), [(g.x, g.y, g.z)])
    h = F.grid_from_file(str(tmp_path / "grid"), g.shape)
    ours = [h.csi[0], h.csi[1], h.csi[2], h.eta[0], h.eta[1], h.eta[2],
            h.zet[0], h.zet[1], h.zet[2], h.aj]
    for n, a in enumerate(ours):
        assert np.array_equal(a, fx.out("metrics", n)), n


DEBUG_FORTRAN = r"""
program dbg
  implicit none
  integer, parameter :: ni = 5, nj = 4, nk = 3, g = 2
  real(8) :: v(1-g:ni+g, 1-g:nj+g, 1-g:nk+g), p(ni, nj, nk)
  integer :: i, j, k
  v = -1.0d0
  do k = 1, nk
  do j = 1, nj
  do i = 1, ni
     v(i,j,k) = i + 10.0d0 * j + 100.0d0 * k
  end do
  end do
  end do
  ! outputD1_real's record (global_debug.F90:49-55): the field without ghost points
  open(43, file='debugfiles/phi.07', form='unformatted')
  write(43) v(1:ni, 1:nj, 1:nk)
  close(43)
  open(44, file='debugfiles/py.07', form='unformatted')
  read(44) p
  close(44)
  print '(es25.16)', sum(p * v(1:ni, 1:nj, 1:nk))
end program dbg
"""


def test_debug_path():
    assert F.debug_path("out/", "phicorrected", 3) == "out/debugfiles/phicorrected.03"
    assert F.debug_path("out/", "phi", 123).endswith("phi.**")


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_debug_dump_fortran_compatibility(tmp_path):
    """A dump written by Fortran (one unformatted record of the nodes without
    ghost layers) reads back, and one written here reads in Fortran."""
    (tmp_path / "debugfiles").mkdir()
    (tmp_path / "dbg.f90").write_text(DEBUG_FORTRAN)
    subprocess.run(["amdflang", "dbg.f90", "-o", "dbg"], cwd=tmp_path, check=True)
    shape = (3, 4, 5)
    kk, jj, ii = np.meshgrid(np.arange(1, 4), np.arange(1, 5), np.arange(1, 6), indexing="ij")
    ref = ii + 10.0 * jj + 100.0 * kk
    rng = np.random.default_rng(7)
    mine = rng.standard_normal(shape)
    padded = np.full((3 + 4, 4 + 4, 5 + 4), -1.0)
    padded[2:-2, 2:-2, 2:-2] = mine
    F.write_debug(F.debug_path(str(tmp_path) + "/", "py", 7), padded, ghost=(2, 2, 2))
    out = subprocess.run(["./dbg"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    assert abs(float(out) - np.sum(mine * ref)) <= 1e-12 * np.sum(np.abs(mine * ref))
    got = F.read_debug(F.debug_path(str(tmp_path) + "/", "phi", 7), shape)
    assert np.array_equal(got, ref)
