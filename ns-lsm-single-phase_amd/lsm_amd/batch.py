"""Batched multi-block level-set advection (the lsm_batch_* C ABI).

A connected batch is the block decomposition of one curvilinear domain, the
reference's MPI decomposition (init.F90:489-1068) held on one GPU: a lattice
of blocks, block (I,J,K) owning global nodes [start_I, start_I + n_I) along
each axis, `ghost` layers on every face filled from the neighbouring blocks
after every RK stage (rhs_exchng3_3d.F90:6-359).  An independent batch holds
unrelated instances without ghost layers.

Batch fields are one flat float64 CUDA tensor of `layout.total` doubles:
block b's array, shape (nk+2g, nj+2g, ni+2g) i fastest with owned nodes at
local indices g .. g+n-1, starts at `layout.offsets[b]`; its rows are
`layout.pitch[b]` doubles apart.  By default rows are padded to a multiple of
16 doubles and every block starts so that its first owned column lies on a
128-byte line (the stage kernels' wavefronts then read whole lines).

multiblock_case() builds the multi-block workload: per-axis block sizes drawn
from U{nmin..nmax}, a per-block sinusoidal perturbation of a uniform grid of
the unit cube (amplitudes U[0, amax]; it vanishes on block faces, so the
blocks join into one continuous grid), the signed distance to random spheres
and a divergence-free random Fourier velocity.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib, inputs
from ._lib import LsmBatchBlock, LsmBatchDesc, LsmError, check

FORMS = {"queue": _lib.LSM_BATCH_QUEUE, "grid": _lib.LSM_BATCH_GRID, "perblock": _lib.LSM_BATCH_PERBLOCK}
FORM_NAMES = {v: k for k, v in FORMS.items()}


@dataclass
class BlockSpec:
    ni: int
    nj: int
    nk: int
    lattice: tuple = (0, 0, 0)
    gstart: tuple = (0, 0, 0)          # global index of the first owned node (connected)
    phys_faces: int = _lib.LSM_FACE_ALL
    btype: tuple = (1, 1, 1, 1, 1, 1)
    big_phi: float = 5.0

    @property
    def owned(self):
        return (self.nk, self.nj, self.ni)


@dataclass
class Layout:
    blocks: List[BlockSpec]
    ghost: int
    connected: bool
    lattice: tuple = (1, 1, 1)
    sizes: tuple = ((), (), ())         # owned sizes of each lattice slab per axis (i, j, k)
    offsets: List[int] = field(default_factory=list)
    total: int = 0
    pitch: List[int] = field(default_factory=list)   # doubles per row of each block's array
    align: int = 16

    def shape(self, b):
        g = self.ghost
        s = self.blocks[b]
        return (s.nk + 2 * g, s.nj + 2 * g, s.ni + 2 * g)

    def alloc_shape(self, b):
        nk, nj, _ = self.shape(b)
        return (nk, nj, self.pitch[b])

    def global_shape(self):
        return tuple(int(sum(self.sizes[a])) for a in (2, 1, 0))   # (Nk, Nj, Ni)

    def cells(self):
        return sum(s.ni * s.nj * s.nk for s in self.blocks)


def _finish(lay: Layout) -> Layout:
    """Offsets and row pitches.  align > 0: rows padded to a multiple of align
    doubles and offset + ghost a multiple of align (first owned column on a
    line boundary); align = 0: dense rows, 256-byte aligned block arrays."""
    off = 0
    a = int(lay.align)
    g = lay.ghost
    lay.offsets = []
    lay.pitch = []
    for b in range(len(lay.blocks)):
        nk, nj, ni = lay.shape(b)
        if a > 0:
            pitch = (ni + a - 1) // a * a
            off += (-(off + g)) % a
        else:
            pitch = ni
            off = (off + 31) // 32 * 32
        lay.offsets.append(off)
        lay.pitch.append(pitch)
        off += pitch * nj * nk
    lay.total = (off + 31) // 32 * 32
    return lay


def lattice_layout(sizes, ghost=3, btype=(1, 1, 1, 1, 1, 1), big_phi=5.0, align=16) -> Layout:
    """Connected layout: sizes = (sizes_i, sizes_j, sizes_k), owned nodes of
    each lattice slab along each axis; align as in _finish."""
    L = tuple(len(s) for s in sizes)
    starts = [np.concatenate([[0], np.cumsum(s)]).astype(int) for s in sizes]
    blocks = []
    for K in range(L[2]):
        for J in range(L[1]):
            for I in range(L[0]):
                f = 0
                if I == 0: f |= 0x01
                if I == L[0] - 1: f |= 0x02
                if J == 0: f |= 0x04
                if J == L[1] - 1: f |= 0x08
                if K == 0: f |= 0x10
                if K == L[2] - 1: f |= 0x20
                blocks.append(BlockSpec(int(sizes[0][I]), int(sizes[1][J]), int(sizes[2][K]), (I, J, K),
                                        (int(starts[0][I]), int(starts[1][J]), int(starts[2][K])), f,
                                        tuple(btype), big_phi))
    return _finish(Layout(blocks, ghost, True, L, tuple(tuple(int(v) for v in s) for s in sizes), align=align))


def independent_layout(shapes, btype=(1, 1, 1, 1, 1, 1), big_phi=5.0, align=16) -> Layout:
    """Independent instances of owned sizes shapes = [(ni, nj, nk), ...]."""
    blocks = [BlockSpec(int(a), int(b), int(c), btype=tuple(btype), big_phi=big_phi) for a, b, c in shapes]
    return _finish(Layout(blocks, 0, False, align=align))


# ---------------------------------------------------------------------------
# the multi-block workload
# ---------------------------------------------------------------------------

def _slab_of(G, start, n):
    """Lattice slab owning global index G (clamped to the domain)."""
    L = np.searchsorted(start, G, side="right") - 1
    return np.clip(L, 0, len(n) - 1)


def block_coordinates(lay: Layout, amps, lo, hi):
    """Node coordinates of global index ranges lo..hi-1 (per axis i, j, k;
    may reach outside the domain): x_a = xi_a + A_o (n_a(o)/(N_a-1)) b with
    xi_a = G_a/(N_a-1), b = prod_a sin(pi t_a), t_a the position of G_a in its
    owner slab, A_o the amplitude of the owner block."""
    N = [int(sum(lay.sizes[a])) for a in range(3)]
    start = [np.concatenate([[0], np.cumsum(lay.sizes[a])]) for a in range(3)]
    nsz = [np.asarray(lay.sizes[a]) for a in range(3)]
    G = [np.arange(lo[a], hi[a]) for a in range(3)]
    xi = [G[a] / (N[a] - 1) for a in range(3)]
    Ls = [_slab_of(G[a], start[a], nsz[a]) for a in range(3)]
    t = [(G[a] - start[a][Ls[a]]) / nsz[a][Ls[a]] for a in range(3)]
    sb = [np.sin(math.pi * t[a]) for a in range(3)]
    bump = sb[2][:, None, None] * sb[1][None, :, None] * sb[0][None, None, :]
    lat = lay.lattice
    owner = Ls[0][None, None, :] + lat[0] * (Ls[1][None, :, None] + lat[1] * Ls[2][:, None, None])
    A = np.asarray(amps)[owner]
    out = []
    for a, ax in ((0, 2), (1, 1), (2, 0)):
        shp = [1, 1, 1]
        shp[ax] = -1
        scale = nsz[a][Ls[a]] / (N[a] - 1)
        out.append(xi[a].reshape(shp) + A * scale.reshape(shp) * bump)
    return [np.ascontiguousarray(np.broadcast_to(c, bump.shape)) for c in out]


def block_grid(lay: Layout, b: int, amps) -> inputs.Grid:
    """Coordinates and metrics of block b's array (ghost layers included).
    Metrics are differences of the coordinates inside the domain, so at owned
    nodes they equal those of the assembled single-block grid; ghost nodes
    outside the domain repeat the nearest inside values."""
    s = lay.blocks[b]
    g = lay.ghost
    N = [int(sum(lay.sizes[a])) for a in range(3)] if lay.connected else [s.ni, s.nj, s.nk]
    gs = s.gstart
    n = (s.ni, s.nj, s.nk)
    lo = [max(gs[a] - g, 0) for a in range(3)]
    hi = [min(gs[a] + n[a] + g, N[a]) for a in range(3)]
    x, y, z = block_coordinates(lay, amps, lo, hi)
    sub = inputs.compute_metrics(inputs.Grid(x, y, z))
    # place into the ghosted array, edge-padding the outside
    pads = [(lo[a] - (gs[a] - g), (gs[a] + n[a] + g) - hi[a]) for a in (2, 1, 0)]

    def pad(f):
        return np.ascontiguousarray(np.pad(f, pads, mode="edge"))
    grid = inputs.Grid(pad(x), pad(y), pad(z))
    grid.aj = pad(sub.aj)
    grid.csi = np.stack([pad(c) for c in sub.csi])
    grid.eta = np.stack([pad(c) for c in sub.eta])
    grid.zet = np.stack([pad(c) for c in sub.zet])
    return grid


def multiblock_case(lattice=(4, 4, 4), nmin=32, nmax=160, ghost=3, amax=0.04, seed=6, nspheres=32,
                    seed_phi=7, seed_vel=4, btype=(1, 1, 1, 1, 1, 1), align=16):
    """The multi-block workload: returns (layout, grids, phi, (u, v, w)) with
    per-block numpy arrays (ghost layers included)."""
    rng = np.random.default_rng(seed)
    sizes = tuple(tuple(int(v) for v in rng.integers(nmin, nmax + 1, size=lattice[a])) for a in range(3))
    amps = rng.uniform(0.0, amax, size=int(np.prod(lattice)))
    lay = lattice_layout(sizes, ghost=ghost, btype=btype, align=align)
    grids = [block_grid(lay, b, amps) for b in range(len(lay.blocks))]
    phi = [inputs.spheres_sdf(gd, nspheres=nspheres, seed=seed_phi, rmin=0.05, rmax=0.15) for gd in grids]
    modes = inputs.fourier_modes(16, 4, seed_vel)
    vel = [inputs.fourier_field(gd, *modes) for gd in grids]
    # unit RMS over the owned nodes of the whole domain
    g = ghost
    ss = 0.0
    cnt = 0
    for (u, v, w) in vel:
        sl = (slice(g, -g or None),) * 3
        ss += float(np.sum(u[sl] ** 2 + v[sl] ** 2 + w[sl] ** 2))
        cnt += u[sl].size
    rms = math.sqrt(ss / cnt)
    vel = [(u / rms, v / rms, w / rms) for (u, v, w) in vel]
    return lay, grids, phi, vel, amps


def assemble(lay: Layout, per_block) -> np.ndarray:
    """The single-block array of a connected layout from its blocks' owned nodes."""
    Nk, Nj, Ni = lay.global_shape()
    out = np.empty((Nk, Nj, Ni))
    g = lay.ghost
    for b, s in enumerate(lay.blocks):
        gi, gj, gk = s.gstart
        out[gk:gk + s.nk, gj:gj + s.nj, gi:gi + s.ni] = per_block[b][g:g + s.nk, g:g + s.nj, g:g + s.ni]
    return out


def owned(lay: Layout, b: int, arr):
    g = lay.ghost
    s = lay.blocks[b]
    return arr[g:g + s.nk, g:g + s.nj, g:g + s.ni]


def cfl_dt(lay: Layout, grids, vel, cfl=0.4):
    """dt = cfl / max over owned nodes of sum_d |U^d| (index spacing 1)."""
    m = 0.0
    for b, (gd, (u, v, w)) in enumerate(zip(grids, vel)):
        U = inputs.contravariant(gd, u, v, w)
        m = max(m, float(np.max(owned(lay, b, np.abs(U[0]) + np.abs(U[1]) + np.abs(U[2])))))
    return cfl / m


# ---------------------------------------------------------------------------
# the device batch
# ---------------------------------------------------------------------------

class Batch:
    """A batch resident on the GPU: metrics and aj as batch arrays, the
    library's batch handle (tile lists, halo tables, launch configuration)."""

    def __init__(self, lay: Layout, grids: Sequence[inputs.Grid], device="cuda"):
        import torch
        self.lay = lay
        self.device = torch.device(device)
        self.aj = self.pack([gd.aj for gd in grids])
        self.csi = [self.pack([gd.csi[c] for gd in grids]) for c in range(3)]
        self.eta = [self.pack([gd.eta[c] for gd in grids]) for c in range(3)]
        self.zet = [self.pack([gd.zet[c] for gd in grids]) for c in range(3)]
        L = _lib.load()
        arr = (LsmBatchBlock * len(lay.blocks))()
        for b, s in enumerate(lay.blocks):
            e = arr[b]
            e.ni, e.nj, e.nk = s.ni, s.nj, s.nk
            e.lattice[:] = list(s.lattice)
            e.phys_faces = s.phys_faces
            e.btype[:] = list(s.btype)
            e.big_phi = s.big_phi
            e.dc = e.de = e.dz = 1.0
            e.offset = lay.offsets[b]
            e.pitch = lay.pitch[b]
        d = LsmBatchDesc()
        d.nblocks = len(lay.blocks)
        d.ghost = lay.ghost
        d.connected = 1 if lay.connected else 0
        d.lattice[:] = list(lay.lattice)
        d.blocks = C.cast(arr, C.POINTER(LsmBatchBlock))
        d.aj = self.aj.data_ptr()
        d.csi[:] = [t.data_ptr() for t in self.csi]
        d.eta[:] = [t.data_ptr() for t in self.eta]
        d.zet[:] = [t.data_ptr() for t in self.zet]
        d.total = lay.total
        self._arr = arr
        h = C.c_void_p()
        check(L.lsm_batch_create(C.byref(d), C.byref(h)), "lsm_batch_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().lsm_batch_destroy(h)
            except Exception:  # noqa: BLE001
                pass
            self.handle = None

    # fields ---------------------------------------------------------------
    def empty(self):
        import torch
        return torch.empty(self.lay.total, dtype=torch.float64, device=self.device)

    def zeros(self):
        import torch
        return torch.zeros(self.lay.total, dtype=torch.float64, device=self.device)

    def pack(self, per_block) -> "torch.Tensor":
        import torch
        flat = np.zeros(self.lay.total)
        for b, a in enumerate(per_block):
            a = np.asarray(a, dtype=np.float64)
            shp = self.lay.shape(b)
            assert a.size == int(np.prod(shp))
            ash = self.lay.alloc_shape(b)
            n = int(np.prod(ash))
            flat[self.lay.offsets[b]:self.lay.offsets[b] + n].reshape(ash)[:, :, :shp[2]] = a.reshape(shp)
        return torch.as_tensor(flat, device=self.device)

    def unpack(self, t) -> List[np.ndarray]:
        flat = t.detach().cpu().numpy()
        out = []
        for b in range(len(self.lay.blocks)):
            shp = self.lay.shape(b)
            ash = self.lay.alloc_shape(b)
            n = int(np.prod(ash))
            out.append(flat[self.lay.offsets[b]:self.lay.offsets[b] + n].reshape(ash)[:, :, :shp[2]].copy())
        return out

    def _check(self, t, name):
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64 or \
                not t.is_contiguous() or t.numel() != self.lay.total:
            raise ValueError(f"{name}: expected a contiguous float64 CUDA tensor of {self.lay.total} elements")
        return C.c_void_p(t.data_ptr())

    # operations -----------------------------------------------------------
    def set_config(self, form="queue", tile_rows=4, kchunk=8, wg_per_cu=0, tile_cols=64):
        check(_lib.load().lsm_batch_set_config(self.handle, FORMS[form] if isinstance(form, str) else int(form),
                                               int(tile_cols), int(tile_rows), int(kchunk), int(wg_per_cu)),
              "lsm_batch_set_config")

    def set_reinit_tile(self, cols=0):
        """Tile columns of the reinitialisation stage kernel (64, 32, 16; 0 =
        the default, 64)."""
        check(_lib.load().lsm_batch_set_reinit_tile(self.handle, int(cols)), "lsm_batch_set_reinit_tile")

    def get_reinit_tile(self):
        return int(_lib.load().lsm_batch_get_reinit_tile(self.handle))

    def get_config(self):
        cfg = (C.c_int32 * 5)()
        check(_lib.load().lsm_batch_get_config(self.handle, cfg), "lsm_batch_get_config")
        return {"form": FORM_NAMES[cfg[0]], "tile_cols": cfg[1], "tile_rows": cfg[2], "kchunk": cfg[3],
                "wg_per_cu": cfg[4]}

    def halo_fill(self, field):
        from .api import _stream
        check(_lib.load().lsm_batch_halo_fill(self.handle, self._check(field, "field"), _stream()),
              "lsm_batch_halo_fill")
        return field

    def contra_velocity(self, u, v, w):
        from .api import _stream
        out = [self.empty() for _ in range(3)]
        check(_lib.load().lsm_batch_contra_velocity(self.handle, self._check(u, "u"), self._check(v, "v"),
                                                    self._check(w, "w"), *[C.c_void_p(t.data_ptr()) for t in out],
                                                    _stream()), "lsm_batch_contra_velocity")
        return out

    def reinit(self, phi0, params, out=None, work=None):
        """reinitialisation_test2 on every block (ghost layers frozen);
        returns (phi_out, iterations per block as an int32 CUDA tensor)."""
        import torch
        from .api import _stream
        if out is None:
            out = self.empty()
        if work is None:
            work = self.workspace(_lib.LSM_OP_REINIT_ENO2)
        iters = torch.zeros(len(self.lay.blocks), dtype=torch.int32, device=self.device)
        p = params.c()
        check(_lib.load().lsm_batch_reinit_sussman_eno2(self.handle, C.byref(p), self._check(phi0, "phi0"),
                                                        self._check(out, "out"), C.c_void_p(iters.data_ptr()),
                                                        C.c_void_p(work.data_ptr()), work.numel(), _stream()),
              "lsm_batch_reinit_sussman_eno2")
        return out, iters

    def bcond(self, field):
        from .api import _stream
        check(_lib.load().lsm_batch_bcond(self.handle, self._check(field, "field"), _stream()), "lsm_batch_bcond")
        return field

    def levelset_step(self, phi_n, ucn, delti_lsm, params=None, out=None, mode="fast", work=None):
        """ad_RKTVD3 (+ reinitialisation_test2, bcond_lsm, halo fill when
        params is given) for every block."""
        from .api import MODES, _stream
        if out is None:
            out = self.empty()
        op = _lib.LSM_OP_LEVELSET_STEP if params is not None else _lib.LSM_OP_RK3
        if work is None:
            work = self.workspace(op)
        p = params.c() if params is not None else None
        check(_lib.load().lsm_batch_levelset_step(self.handle, C.byref(p) if p is not None else None,
                                                  self._check(out, "out"), self._check(phi_n, "phi_n"),
                                                  *[self._check(t, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                                  C.c_void_p(work.data_ptr()), work.numel(), _stream()),
              "lsm_batch_levelset_step")
        return out

    def workspace(self, op=_lib.LSM_OP_RK3):
        import torch
        n = _lib.load().lsm_batch_workspace_bytes(self.handle, op)
        if n == 0:
            raise LsmError("lsm_batch_workspace_bytes failed")
        return torch.zeros(n, dtype=torch.uint8, device=self.device)

    def rk3_step(self, phi_n, ucn, delti_lsm, out=None, mode="fast", work=None):
        from .api import MODES, _stream
        if out is None:
            out = self.empty()
        if work is None:
            work = self.workspace()
        check(_lib.load().lsm_batch_rk3_step(self.handle, self._check(out, "out"), self._check(phi_n, "phi_n"),
                                             *[self._check(t, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                             C.c_void_p(work.data_ptr()), work.numel(), _stream()),
              "lsm_batch_rk3_step")
        return out

    def autotune(self, phi_n, ucn, delti_lsm, mode="fast", reps=3, force=False, out=None, work=None):
        """Picks the fastest stage launch configuration for this batch's
        shapes (lsm_batch_autotune); returns (config, ms per step or None if
        it came from the cache).  `out` is overwritten."""
        from .api import MODES, _stream
        if out is None:
            out = self.empty()
        if work is None:
            work = self.workspace()
        ms = C.c_double()
        check(_lib.load().lsm_batch_autotune(self.handle, self._check(out, "out"), self._check(phi_n, "phi_n"),
                                             *[self._check(t, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                             C.c_void_p(work.data_ptr()), work.numel(), int(reps), int(bool(force)),
                                             C.byref(ms), _stream()), "lsm_batch_autotune")
        return self.get_config(), (ms.value if ms.value >= 0 else None)

    def autotune_reinit(self, phi_n, ucn, delti_lsm, params, mode="fast", reps=3, out=None, work=None,
                        widths=(64, 32, 16)):
        """Times one full step (levelset_step with params) per reinitialisation
        tile width, keeps the fastest; returns (width, {width: ms per step}).
        `out` is overwritten."""
        import torch
        if out is None:
            out = self.empty()
        if work is None:
            work = self.workspace(_lib.LSM_OP_LEVELSET_STEP)
        times = {}
        for w in widths:
            self.set_reinit_tile(w)
            self.levelset_step(phi_n, ucn, delti_lsm, params, mode=mode, out=out, work=work)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.levelset_step(phi_n, ucn, delti_lsm, params, mode=mode, out=out, work=work)
            e1.record()
            torch.cuda.synchronize()
            times[w] = e0.elapsed_time(e1) / reps
        best = min(times, key=times.get)
        self.set_reinit_tile(best)
        return best, times

    def graph_capture(self, phi_np1, phi_n, ucn, delti_lsm, params=None, mode="fast", work=None):
        """Records one levelset_step with these tensors into a HIP graph kept
        by the batch (the tensors must stay alive while the graph is used)."""
        from .api import MODES
        op = _lib.LSM_OP_LEVELSET_STEP if params is not None else _lib.LSM_OP_RK3
        if work is None:
            work = self.workspace(op)
        p = params.c() if params is not None else None
        self._graph_keep = (phi_np1, phi_n, tuple(ucn), work, p)
        check(_lib.load().lsm_batch_graph_capture(self.handle, C.byref(p) if p is not None else None,
                                                  self._check(phi_np1, "phi_np1"), self._check(phi_n, "phi_n"),
                                                  *[self._check(t, "ucn") for t in ucn], float(delti_lsm), MODES[mode],
                                                  C.c_void_p(work.data_ptr()), work.numel()),
              "lsm_batch_graph_capture")

    def graph_launch(self):
        from .api import _stream
        check(_lib.load().lsm_batch_graph_launch(self.handle, _stream()), "lsm_batch_graph_launch")
