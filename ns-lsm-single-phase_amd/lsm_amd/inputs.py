"""Synthetic inputs for the level-set transport path: curvilinear grids, their
metrics, initial level-set fields and frozen velocity fields.

Array convention: every field is a C-contiguous ``float64`` array of shape
``(nk, nj, ni)``, so the flat memory is i-fastest, the same order as the
reference's Fortran arrays ``phi(il:iu, jl:ju, kl:ku)``.  Metric vectors are
kept as separate components (structure of arrays): ``csi[0]`` is xi_x,
``csi[1]`` xi_y, ``csi[2]`` xi_z, and likewise for ``eta`` and ``zet``.

Input generation runs on the host; it is not part of the accelerated path.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np


@dataclass
class Grid:
    """One curvilinear block: node coordinates plus the metric terms the
    transport kernels read (``csi``, ``eta``, ``zet`` = grad xi, grad eta,
    grad zeta, and ``aj`` = 1/det(dx/dxi), the reference's ``aj``)."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    csi: np.ndarray = field(default=None)
    eta: np.ndarray = field(default=None)
    zet: np.ndarray = field(default=None)
    aj: np.ndarray = field(default=None)
    dc: float = 1.0
    de: float = 1.0
    dz: float = 1.0

    @property
    def shape(self):
        return self.x.shape

    @property
    def ni(self):
        return self.x.shape[2]

    @property
    def nj(self):
        return self.x.shape[1]

    @property
    def nk(self):
        return self.x.shape[0]


def _dxi(f, axis, d):
    """d f / d xi along ``axis`` as mg_metrics.F90:53-170 computes it for a
    block whose faces are all physical: central difference inside
    (0.5*d*(f[+1]-f[-1])), first-order one-sided at the two end nodes."""
    out = np.empty_like(f)
    half = 0.5 * d
    sl = [slice(None)] * 3

    def s(a, b):
        q = list(sl)
        q[axis] = slice(a, b)
        return tuple(q)

    n = f.shape[axis]
    out[s(1, n - 1)] = half * (f[s(2, n)] - f[s(0, n - 2)])
    out[s(0, 1)] = d * (f[s(1, 2)] - f[s(0, 1)])
    out[s(n - 1, n)] = d * (f[s(n - 1, n)] - f[s(n - 2, n - 1)])
    return out


def compute_metrics(g: Grid) -> Grid:
    """Metrics of the mapping, following mg_metrics.F90:53-262 with the
    operation order of that file (gj, aj = 1/gj, csi/eta/zet = aj*cofactors).
    Axis 2 is i (xi), axis 1 is j (eta), axis 0 is k (zeta)."""
    x, y, z = g.x, g.y, g.z
    xc, yc, zc = (_dxi(a, 2, g.dc) for a in (x, y, z))
    xe, ye, ze = (_dxi(a, 1, g.de) for a in (x, y, z))
    xz, yz, zz = (_dxi(a, 0, g.dz) for a in (x, y, z))
    gj = (xc * (ye * zz - yz * ze)
          + xe * (yz * zc - yc * zz)) + xz * (yc * ze - ye * zc)
    if np.any(gj == 0.0):
        raise ValueError("zero Jacobian in grid mapping")
    aj = 1.0 / gj
    g.aj = aj
    g.csi = np.stack([aj * (ye * zz - yz * ze),
                      aj * (xz * ze - xe * zz),
                      aj * (xe * yz - xz * ye)])
    g.eta = np.stack([aj * (yz * zc - yc * zz),
                      aj * (xc * zz - xz * zc),
                      aj * (xz * yc - xc * yz)])
    g.zet = np.stack([aj * (yc * ze - ye * zc),
                      aj * (xe * zc - xc * ze),
                      aj * (xc * ye - xe * yc)])
    return g


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------

def _two_sided_tanh(n, length, zc, beta):
    """Nodes on [0, length] clustered toward the interior point ``zc``: two
    tanh pieces meet at zc with equal slope, each densest at zc."""
    s = np.linspace(0.0, 1.0, n)
    sc = zc / length                       # equal slopes at the junction
    tb = math.tanh(beta)
    z = np.where(s <= sc,
                 zc * np.tanh(beta * s / sc) / tb,
                 length - (length - zc) * np.tanh(beta * (1.0 - s) / (1.0 - sc)) / tb)
    return z


def channel_grid(ni=96, nj=24, nk=48, length=20.0, width=2.0, depth=1.5,
                 d=1.0, beta=2.0) -> Grid:
    """Solitary-wave channel: x in [0, length] and y in [0, width] uniform,
    z in [0, depth] tanh-stretched (beta) so nodes cluster toward z = d."""
    xs = np.linspace(0.0, length, ni)
    ys = np.linspace(0.0, width, nj)
    zs = _two_sided_tanh(nk, depth, d, beta)
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    return compute_metrics(Grid(x.copy(), y.copy(), z.copy()))


def sinusoidal_cube_grid(n=256, amp=0.03, seed=2, nj=None, nk=None) -> Grid:
    """Unit cube with x_a = xi_a + amp*sin(2 pi xi_b + t_{2a-1})*sin(2 pi xi_c + t_{2a})
    for cyclic (a, b, c); the six phases are U[0, 2 pi) from ``seed``."""
    nj = n if nj is None else nj
    nk = n if nk is None else nk
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, 2.0 * math.pi, 6)
    xi = np.linspace(0.0, 1.0, n)
    et = np.linspace(0.0, 1.0, nj)
    ze = np.linspace(0.0, 1.0, nk)
    Z, E, X = np.meshgrid(ze, et, xi, indexing="ij")
    tp = 2.0 * math.pi
    x = X + amp * np.sin(tp * E + th[0]) * np.sin(tp * Z + th[1])
    y = E + amp * np.sin(tp * Z + th[2]) * np.sin(tp * X + th[3])
    z = Z + amp * np.sin(tp * X + th[4]) * np.sin(tp * E + th[5])
    return compute_metrics(Grid(x, y, z))


def uniform_grid(ni, nj, nk, h=1.0) -> Grid:
    """Cartesian grid with spacing h (metrics have a closed form: xi_x = 1/h,
    aj = 1/h^3)."""
    xs = np.arange(ni) * h
    ys = np.arange(nj) * h
    zs = np.arange(nk) * h
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    return compute_metrics(Grid(x.copy(), y.copy(), z.copy()))


# ---------------------------------------------------------------------------
# level-set fields
# ---------------------------------------------------------------------------

def solitary_wave(x, H=0.3, d=1.0, x0=5.0):
    """First-order solitary wave elevation eta and its slope."""
    k = math.sqrt(3.0 * H / (4.0 * d ** 3))
    s = 1.0 / np.cosh(k * (x - x0))
    eta = H * s * s
    deta = -2.0 * H * k * s * s * np.tanh(k * (x - x0))
    return eta, deta


def solitary_phi(g: Grid, H=0.3, d=1.0, x0=5.0):
    """phi = z - (d + H sech^2(k (x - x0)))."""
    eta, _ = solitary_wave(g.x, H, d, x0)
    return g.z - (d + eta)


def spheres_sdf(g: Grid, nspheres=8, seed=3, cmin=0.2, cmax=0.8,
                rmin=0.05, rmax=0.15):
    """Signed distance to a union of spheres, positive inside:
    phi = max_s (r_s - |x - c_s|)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(cmin, cmax, (nspheres, 3))
    r = rng.uniform(rmin, rmax, nspheres)
    phi = np.full(g.x.shape, -np.inf)
    for s in range(nspheres):
        dist = np.sqrt((g.x - c[s, 0]) ** 2 + (g.y - c[s, 1]) ** 2 + (g.z - c[s, 2]) ** 2)
        np.maximum(phi, r[s] - dist, out=phi)
    return phi


# ---------------------------------------------------------------------------
# velocity fields (frozen)
# ---------------------------------------------------------------------------

def solitary_velocity(g: Grid, H=0.3, d=1.0, x0=5.0, gacc=9.81):
    """u = c eta/d, v = 0, w from continuity with w(z=0) = 0:
    w = -(c/d) eta'(x) z, c = sqrt(g (d + H))."""
    c = math.sqrt(gacc * (d + H))
    eta, deta = solitary_wave(g.x, H, d, x0)
    u = c * eta / d
    v = np.zeros_like(u)
    w = -(c / d) * deta * g.z
    return u, v, w


def fourier_modes(nmodes=16, kmax=4, seed=4):
    """The random modes of fourier_velocity: wavevectors k_m (nmodes x 3,
    integer-valued), unit directions e_m normal to k_m, amplitudes a_m ~ N(0,1)
    and phases psi_m ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    cand = np.array([(a, b, c) for a in range(-kmax, kmax + 1)
                     for b in range(-kmax, kmax + 1)
                     for c in range(-kmax, kmax + 1)
                     if 1 <= a * a + b * b + c * c <= kmax * kmax], dtype=np.float64)
    pick = rng.choice(len(cand), nmodes, replace=False)
    ks = cand[pick]
    es = np.zeros((nmodes, 3))
    amps = np.zeros(nmodes)
    psis = np.zeros(nmodes)
    for m in range(nmodes):
        k = ks[m]
        r = rng.normal(size=3)
        e = r - (r @ k) / (k @ k) * k
        es[m] = e / np.linalg.norm(e)
        amps[m] = rng.normal()
        psis[m] = rng.uniform(0.0, 2.0 * math.pi)
    return ks, es, amps, psis


def fourier_field(g: Grid, ks, es, amps, psis):
    """u = sum_m amps_m e_m cos(2 pi k_m . x + psi_m) at the nodes of g."""
    u = np.zeros(g.x.shape)
    v = np.zeros(g.x.shape)
    w = np.zeros(g.x.shape)
    for m in range(len(amps)):
        k, e = ks[m], es[m]
        ph = np.cos(2.0 * math.pi * (k[0] * g.x + k[1] * g.y + k[2] * g.z) + psis[m])
        u += amps[m] * e[0] * ph
        v += amps[m] * e[1] * ph
        w += amps[m] * e[2] * ph
    return u, v, w


def fourier_velocity(g: Grid, nmodes=16, kmax=4, seed=4, return_modes=False):
    """Divergence-free random Fourier field u = sum_m a_m e_m cos(2 pi k_m.x + psi_m),
    k_m in Z^3 with 1 <= |k| <= kmax, e_m a unit vector normal to k_m,
    a ~ N(0,1), psi ~ U[0, 2 pi); scaled to unit RMS over the nodes.  With
    return_modes the modes are returned too, amplitudes already divided by the
    RMS, so fourier_field(g, *modes) gives the same field."""
    ks, es, amps, psis = fourier_modes(nmodes, kmax, seed)
    u, v, w = fourier_field(g, ks, es, amps, psis)
    rms = math.sqrt(float(np.mean(u * u + v * v + w * w)))
    out = (u / rms, v / rms, w / rms)
    if return_modes:
        return out, (ks, es, amps / rms, psis)
    return out


def contravariant(g: Grid, u, v, w):
    """Contravariant velocity components U^j = grad xi^j . u (index space, the
    reference's aj*ucn_j)."""
    U1 = g.csi[0] * u + g.csi[1] * v + g.csi[2] * w
    U2 = g.eta[0] * u + g.eta[1] * v + g.eta[2] * w
    U3 = g.zet[0] * u + g.zet[1] * v + g.zet[2] * w
    return U1, U2, U3


def cfl_dt(g: Grid, u, v, w, cfl=0.4):
    """dt = cfl / max_ijk sum_d |U^d| (index spacing 1, dc = de = dz = 1)."""
    U1, U2, U3 = contravariant(g, u, v, w)
    return cfl / float(np.max(np.abs(U1) + np.abs(U2) + np.abs(U3)))


def spacing_bounds(g: Grid):
    """(min, max) node spacing along grid lines."""
    hs = []
    for ax in range(3):
        dx = np.diff(g.x, axis=ax)
        dy = np.diff(g.y, axis=ax)
        dz = np.diff(g.z, axis=ax)
        hs.append(np.sqrt(dx * dx + dy * dy + dz * dz))
    return (float(min(h.min() for h in hs)), float(max(h.max() for h in hs)))
