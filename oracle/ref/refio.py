"""Runs the compiled reference Fortran harness (oracle/_ref/lsm_ref_harness).

Test infrastructure only: used to write golden vectors and to time the
reference on the host.  The input/output layout is documented in
lsm_ref_harness.F90.  The same driver runs oracle/_ref/lsm_shim_harness, the
harness built with the GPU drop-in routines in place of the reference's
(`make -C oracle/ref shim`), for the drop-in parity tests.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "..", "_ref", "lsm_ref_harness")
SHIM_HARNESS = os.path.join(HERE, "..", "_ref", "lsm_shim_harness")
# the same harness with levelsetmethod itself replaced by the whole-call drop-in
RESIDENT_HARNESS = os.path.join(HERE, "..", "_ref", "lsm_resident_harness")

MODES = {
    "advect": 1, "weno_rhs": 2, "bcond": 3, "contra": 4, "normgrad": 5,
    "reinit_eno2": 6, "sign_ini": 7, "rh_rn": 8, "lambda_rn": 9,
    "reinit_godunov": 10, "volume": 11, "levelset": 12, "metrics": 13,
    "phi_gradient": 14, "calc_h": 15, "h_gradient": 16, "upwind_rhs": 17,
}


def available() -> bool:
    return os.path.exists(HARNESS)


def _aos(c):
    # (3, nk, nj, ni) SoA -> Fortran csi(1:3, i, j, k) memory order
    return np.ascontiguousarray(np.moveaxis(np.asarray(c), 0, -1))


def run(mode, grid, phi, phi0=None, aux=None, mask=None, u=None, v=None, w=None,
        nsteps=1, btype=(1, 1, 1, 1, 1, 1), order_ad=1, numiter_lsm=1,
        rn_itermax=4, order_reinit_bc=1, eno_bc=False, m_sphi0=1, m_ti=1,
        hybrid=False, call_reinit=True, rn_freq=1, ntime0=1,
        sussman_method=0, call_levelset=True, delti_lsm=1e-3, epslsm=0.05,
        rn_timestep=1e-3, eps_reinit=1e-5, big_phi=5.0, lambda_dt=1e-3,
        timing=False, workdir=None, nprocs=1, harness=None, blanking=None):
    """Run one harness mode; returns (list of outputs, seconds or None, extra).
    harness: path of the executable (default: the reference harness)."""
    nk, nj, ni = grid.shape
    ip = np.zeros(64, dtype="<i4")
    ip[0] = MODES[mode] if isinstance(mode, str) else int(mode)
    ip[1:4] = (ni, nj, nk)
    ip[4] = nsteps
    ip[5:11] = btype
    ip[11] = order_ad
    ip[12] = numiter_lsm
    ip[13] = rn_itermax
    ip[14] = order_reinit_bc
    ip[15] = int(bool(eno_bc))
    ip[16] = m_sphi0
    ip[17] = m_ti
    ip[18] = int(bool(hybrid))
    ip[19] = int(bool(call_reinit))
    ip[20] = rn_freq
    ip[21] = ntime0
    ip[22] = sussman_method
    ip[23] = int(bool(call_levelset))
    ip[24] = int(bool(timing))
    if blanking is not None:
        # one blanking box: ((ia, ib, ja, jb, ka, kb) 0-based, blktype(1:6))
        box, bt = blanking
        ip[30] = 1
        ip[31:37] = [int(v) + 1 for v in box]
        ip[37:43] = list(bt) + [0] * (6 - len(bt))
    rp = np.zeros(32, dtype="<f8")
    rp[0:3] = (grid.dc, grid.de, grid.dz)
    rp[3] = delti_lsm
    rp[4] = epslsm
    rp[5] = rn_timestep
    rp[6] = eps_reinit
    rp[7] = big_phi
    rp[8] = lambda_dt

    zeros = np.zeros(grid.shape)
    u = zeros if u is None else u
    v = zeros if v is None else v
    w = zeros if w is None else w
    q = np.stack([zeros, u, v, w], axis=-1)
    phi0 = zeros if phi0 is None else phi0
    aux = zeros if aux is None else aux
    mask = np.ones(grid.shape, dtype="<i4") if mask is None else mask

    own_dir = workdir is None
    wd = tempfile.mkdtemp(prefix="lsmref_") if own_dir else workdir
    os.makedirs(os.path.join(wd, "debugfiles"), exist_ok=True)
    fin = os.path.join(wd, "in.bin")
    fout = os.path.join(wd, "out.bin")
    with open(fin, "wb") as f:
        f.write(ip.tobytes())
        f.write(rp.tobytes())
        for a in (grid.x, grid.y, grid.z):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for c in (grid.csi, grid.eta, grid.zet):
            f.write(_aos(c).astype("<f8").tobytes())
        f.write(np.ascontiguousarray(grid.aj, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(q, dtype="<f8").tobytes())
        for a in (phi, phi0, aux):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(mask, dtype="<i4").tobytes())

    env = dict(os.environ)
    cmd = [os.path.abspath(harness or HARNESS), fin, fout]
    if nprocs > 1:
        cmd = ["/opt/conda/bin/mpiexec", "-n", str(nprocs)] + cmd

    def big_stack():
        # the reference keeps full-size work arrays as automatic (stack)
        # arrays (ad_RKTVD3.F90:9, reinitialisation_test2.F90:21-26): run it
        # with the stack limit raised to the hard limit ("ulimit -s unlimited")
        import resource
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    r = subprocess.run(cmd, cwd=wd, env=env, capture_output=True, text=True, preexec_fn=big_stack)
    if r.returncode != 0:
        raise RuntimeError(f"reference harness failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    secs = None
    for line in r.stdout.splitlines():
        if line.startswith("REF_SECONDS"):
            t = float(line.split()[1])
            secs = t if secs is None else max(secs, t)   # slowest rank
    outs = []
    raw = open(fout, "rb").read()
    off = 0
    nout = int(np.frombuffer(raw, "<i4", 1, off)[0]); off += 4
    for _ in range(nout):
        kind, n = np.frombuffer(raw, "<i4", 2, off); off += 8
        if kind == 0:
            a = np.frombuffer(raw, "<f8", n, off).copy(); off += 8 * n
        else:
            a = np.frombuffer(raw, "<i4", n, off).copy(); off += 4 * n
        if n == ni * nj * nk:
            a = a.reshape(grid.shape)
        outs.append(a)
    extra = {}
    gv = os.path.join(wd, "GlobalVolume.txt")
    if os.path.exists(gv):
        extra["global_volume"] = np.array([float(t) for t in open(gv).read().split()])
    if own_dir:
        import shutil
        shutil.rmtree(wd, ignore_errors=True)
    return outs, secs, extra
