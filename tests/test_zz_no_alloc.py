"""lsm_water_volume_noalloc: the water volume (TotalWaterVolume.F90:1-297)
with its scratch in a fixed device array of the library instead of a
workspace.  Checked against lsm_water_volume (bit for bit: the same kernels
on the same inputs) on two calls in a row, the second on a smaller block, so
a result cannot come from what the previous call left in the shared array."""
import ctypes as C

import numpy as np
import pytest

from lsm_amd import _lib, inputs


def test_noalloc_rejects_blocks_beyond_its_scratch():
    """Host-side size check, before any device work: the pointers are never
    dereferenced."""
    L = _lib.load()
    b = _lib.LsmBlock()
    b.il, b.iu, b.jl, b.ju, b.kl, b.ku = 1, 400, 1, 400, 1, 400
    b.phys_faces = _lib.LSM_FACE_ALL
    b.btype[:] = [1] * 6
    b.dc = b.de = b.dz = 1.0
    b.big_phi = 5.0
    b.x = b.y = b.z = 0x1000
    fake = C.c_void_p(0x1000)
    rc = L.lsm_water_volume_noalloc(C.byref(b), fake, fake, fake, None)
    assert rc == _lib.LSM_ERR_WORKSPACE, rc
    assert b"scratch" in L.lsm_last_error()
    rc = L.lsm_water_volume_noalloc(C.byref(b), fake, None, None, None)
    assert rc == _lib.LSM_ERR_ARG


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry_point():
    import torch
    from lsm_amd import api
    L = _lib.load()
    for n, seed in ((48, 401), (22, 402)):
        print(f"[no-alloc] block {n}x{n - 3}x{n - 6} seed {seed}", flush=True)
        g = inputs.sinusoidal_cube_grid(n, amp=0.03, seed=seed, nj=n - 3, nk=n - 6)
        phi = torch.as_tensor(inputs.spheres_sdf(g, nspheres=6, seed=seed + 1, rmin=0.1, rmax=0.3), device="cuda")
        blk = api.Block(g)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        ws = blk.workspace(_lib.LSM_OP_WATER_VOLUME)
        v_ref = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
        m_ref = torch.full(blk.shape, -5, dtype=torch.int32, device="cuda")
        _lib.check(L.lsm_water_volume(blk.ref, C.c_void_p(phi.data_ptr()), C.c_void_p(v_ref.data_ptr()),
                                      C.c_void_p(m_ref.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel(), st),
                   "lsm_water_volume")
        v = torch.full((3,), -2.0, dtype=torch.float64, device="cuda")
        m = torch.full(blk.shape, -7, dtype=torch.int32, device="cuda")
        _lib.check(L.lsm_water_volume_noalloc(blk.ref, C.c_void_p(phi.data_ptr()), C.c_void_p(v.data_ptr()),
                                              C.c_void_p(m.data_ptr()), st), "lsm_water_volume_noalloc")
        v2 = torch.full((3,), -3.0, dtype=torch.float64, device="cuda")
        _lib.check(L.lsm_water_volume_noalloc(blk.ref, C.c_void_p(phi.data_ptr()), C.c_void_p(v2.data_ptr()),
                                              None, st), "lsm_water_volume_noalloc (no mask)")
        torch.cuda.synchronize()
        assert torch.equal(v, v_ref) and torch.equal(v2, v_ref), (v, v_ref)
        assert torch.equal(m, m_ref)
        assert float(v_ref.sum()) > 0.0 and int(m_ref.sum()) > 0
