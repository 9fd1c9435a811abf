// Geometry and the |grad d| node formula shared by the reinitialisation
// kernels and their adjoint.
#pragma once

#include "lsm_internal.hpp"

namespace lsm {
namespace rn {

struct Geo {
    int ni, nj, nk;
    long sj, sk;
    double dd[3];
    const double *m[9];   // csi0..2, eta0..2, zet0..2
};

inline Geo make_geo(const Blk &b)
{
    Geo g;
    g.ni = b.e.ni; g.nj = b.e.nj; g.nk = b.e.nk;
    g.sj = b.e.sj; g.sk = b.e.sk;
    g.dd[0] = b.dc; g.dd[1] = b.de; g.dd[2] = b.dz;
    g.m[0] = b.csi0; g.m[1] = b.csi1; g.m[2] = b.csi2;
    g.m[3] = b.eta0; g.m[4] = b.eta1; g.m[5] = b.eta2;
    g.m[6] = b.zet0; g.m[7] = b.zet1; g.m[8] = b.zet2;
    return g;
}

__device__ __forceinline__ void node_of(const Geo &g, long l, int &i, int &j, int &k)
{
    i = (int)(l % g.ni);
    const long r = l / g.ni;
    j = (int)(r % g.nj);
    k = (int)(r / g.nj);
}

// The metric tensor G_ab = grad xi_a . grad xi_b of one node from its nine
// metrics m (csi, eta, zet: the gradients of xi, eta, zeta), stored in the
// order G_xx, G_xe, G_xz, G_ee, G_ez, G_zz.  |grad d|^2 = g^T G g for the
// computational derivatives g: 6 fields per node instead of 9.
__device__ __forceinline__ void metric_tensor(const double m[9], double G[6])
{
    G[0] = (m[0] * m[0] + m[1] * m[1]) + m[2] * m[2];
    G[1] = (m[0] * m[3] + m[1] * m[4]) + m[2] * m[5];
    G[2] = (m[0] * m[6] + m[1] * m[7]) + m[2] * m[8];
    G[3] = (m[3] * m[3] + m[4] * m[4]) + m[5] * m[5];
    G[4] = (m[3] * m[6] + m[4] * m[7]) + m[5] * m[8];
    G[5] = (m[6] * m[6] + m[7] * m[7]) + m[8] * m[8];
}

// |grad d| from the tensor and the computational derivatives g0, g1, g2;
// t = G g (t_a / |grad d| = d|grad d| / d g_a).  The quadratic form of a
// positive definite G is clamped at 0 against rounding.
__device__ __forceinline__ double tensor_norm(const double G[6], double g0, double g1, double g2, double t[3])
{
    t[0] = (G[0] * g0 + G[1] * g1) + G[2] * g2;
    t[1] = (G[1] * g0 + G[3] * g1) + G[4] * g2;
    t[2] = (G[2] * g0 + G[4] * g1) + G[5] * g2;
    return sqrt(fmax((g0 * t[0] + g1 * t[1]) + g2 * t[2], 0.0));
}

// Derivatives of d along xi, eta, zeta at node (i,j,k) (not a vertex of a
// six-face block) with the stencils of calc_NormGrad_d.F90:39-2446: ENO2 with
// upwind sign p0 where the direction is not normal to a physical face, the
// one-sided formula of order order_bc on the face, and the fixed second-order
// xi formula on the edge i = 1, k = 1 (calc_NormGrad_d.F90:1281-1283; the
// same in calc_phi_gradient_ENO2.F90:1413-1415).
// blo / bhi (bit a: axis a) force the interior ENO2 stencil one-sided as the
// physical faces do (the blanking biasing of the consumers of phi).
__device__ __forceinline__ void curv_derivs(const Geo &g, const double *__restrict__ d, double p0,
                                            int order_bc, bool eno_bc, int i, int j, int k, long l,
                                            const bool on[3], double gr[3], int blo = 0, int bhi = 0)
{
    const int n[3] = {g.ni, g.nj, g.nk};
    const int idx[3] = {i, j, k};
    const long st[3] = {1, g.sj, g.sk};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long s = st[a];
        if (on[a]) {
            const int dir = (idx[a] == 0) ? 1 : -1;
            const long ss = s * dir;
            if (a == 0 && i == 0 && k == 0 && !on[1]) {
                const double dc2 = 0.5 * g.dd[0];
                gr[a] = dc2 * ((-(1.0 * d[l + 2]) + 4.0 * d[l + 1]) - 3.0 * d[l]);
            } else {
                gr[a] = g.dd[a] * biased(d[l], d[l + ss], d[l + 2 * ss], d[l + 3 * ss], order_bc, dir, eno_bc);
            }
        } else {
            const bool lo = (idx[a] == 1) || ((blo >> a) & 1), hi = (idx[a] == n[a] - 2) || ((bhi >> a) & 1);
            const double fLL = lo ? 0.0 : d[l - 2 * s];
            const double fRR = hi ? 0.0 : d[l + 2 * s];
            gr[a] = g.dd[a] * eno2(p0, fLL, d[l - s], d[l], d[l + s], fRR, lo, hi);
        }
    }
}

// |grad d| at node (i,j,k): calc_NormGrad_d.F90:39-2446 for a block whose six
// faces are physical.  phi0s is phizero at the last interior node, which the
// i = 1 / i = im face blocks use for their ENO2 upwind sign (those blocks never
// assign phi0, calc_NormGrad_d.F90:187-367).
__device__ inline double normgrad_node(const Geo &g, const double *__restrict__ d,
                                const double *__restrict__ phi0, double phi0s, int order_bc,
                                bool eno_bc, int i, int j, int k, long l)
{
    bool on[3];
    on[0] = (i == 0 || i == g.ni - 1);
    on[1] = (j == 0 || j == g.nj - 1);
    on[2] = (k == 0 || k == g.nk - 1);
    const double p0 = (on[0] && !on[1] && !on[2]) ? phi0s : phi0[l];
    double gr[3];
    curv_derivs(g, d, p0, order_bc, eno_bc, i, j, k, l, on, gr);
    const double gx = (g.m[0][l] * gr[0] + g.m[3][l] * gr[1]) + g.m[6][l] * gr[2];
    const double gy = (g.m[1][l] * gr[0] + g.m[4][l] * gr[1]) + g.m[7][l] * gr[2];
    const double gz = (g.m[2][l] * gr[0] + g.m[5][l] * gr[1]) + g.m[8][l] * gr[2];
    return sqrt((gx * gx + gy * gy) + gz * gz);
}

// normgrad_node with |grad d|^2 = g^T G g from the metric tensor gt of the
// node (component c at gt + c * gs), the form of the reinitialisation stages
__device__ inline double normgrad_node_g(const Geo &g, const double *__restrict__ d,
                                         const double *__restrict__ phi0, const double *gt, long gs, double phi0s,
                                         int order_bc, bool eno_bc, int i, int j, int k, long l)
{
    bool on[3];
    on[0] = (i == 0 || i == g.ni - 1);
    on[1] = (j == 0 || j == g.nj - 1);
    on[2] = (k == 0 || k == g.nk - 1);
    const double p0 = (on[0] && !on[1] && !on[2]) ? phi0s : phi0[l];
    double gr[3], G[6], t[3];
    curv_derivs(g, d, p0, order_bc, eno_bc, i, j, k, l, on, gr);
#pragma unroll
    for (int c = 0; c < 6; ++c) G[c] = gt[c * gs + l];
    return tensor_norm(G, gr[0], gr[1], gr[2], t);
}

// branch codes of the adjoint: bits 0-3, 4-7, 8-11 the difference formula of
// each direction, kLive: the node's update depends on its stencil, kHasT: it
// depends on s(phi0) (inside the interface band)
constexpr uint16_t kLive = 0x8000;
constexpr uint16_t kHasT = 0x4000;

// |grad d| at node (i,j,k) (the formula of normgrad_node_g), the direction
// weights w_d = (grad d . grad xi_d / |grad d|) x spacing and the branch codes
// of the three difference formulas (4 bits each)
__device__ inline double normgrad_branch(const Geo &g, const double *__restrict__ d, const double *__restrict__ phi0,
                                  const double *gt, long gs, double phi0s, int order_bc, bool eno_bc, int i, int j,
                                  int k, long l, double w[3], int &code)
{
    const int n[3] = {g.ni, g.nj, g.nk};
    const int idx[3] = {i, j, k};
    const long st[3] = {1, g.sj, g.sk};
    bool on[3];
    for (int a = 0; a < 3; ++a) on[a] = (idx[a] == 0 || idx[a] == n[a] - 1);
    const double p0 = (on[0] && !on[1] && !on[2]) ? phi0s : phi0[l];
    double gr[3], scale[3];
    code = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long s = st[a];
        int br = 0;
        scale[a] = g.dd[a];
        if (on[a]) {
            const int dir = (idx[a] == 0) ? 1 : -1;
            const long ss = s * dir;
            if (a == 0 && i == 0 && k == 0 && !on[1]) {
                // edge i = 1, k = 1 (calc_NormGrad_d.F90:1281-1283)
                scale[a] = 0.5 * g.dd[0];
                gr[a] = scale[a] * ((-(1.0 * d[l + 2]) + 4.0 * d[l + 1]) - 3.0 * d[l]);
            } else {
                const double f0 = d[l], f1 = d[l + ss], f2 = d[l + 2 * ss], f3 = d[l + 3 * ss];
                gr[a] = g.dd[a] * biased(f0, f1, f2, f3, order_bc, dir, eno_bc);
                if (order_bc == 2 && eno_bc) br = eno2_boundary_branch(f0, f1, f2, dir);
            }
        } else {
            const bool lo = (idx[a] == 1), hi = (idx[a] == n[a] - 2);
            const double fLL = lo ? 0.0 : d[l - 2 * s];
            const double fRR = hi ? 0.0 : d[l + 2 * s];
            const double fL = d[l - s], fC = d[l], fR = d[l + s];
            gr[a] = g.dd[a] * eno2(p0, fLL, fL, fC, fR, fRR, lo, hi);
            br = eno2_branch(p0, fLL, fL, fC, fR, fRR, lo, hi);
        }
        code |= br << (4 * a);
    }
    double G[6], t[3];
#pragma unroll
    for (int c = 0; c < 6; ++c) G[c] = gt[c * gs + l];
    const double ng = tensor_norm(G, gr[0], gr[1], gr[2], t);   // t_a = grad d . grad xi_a
    const double inv = ng > 0.0 ? 1.0 / ng : 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = t[a] * inv * scale[a];
    return ng;
}

}  // namespace rn
}  // namespace lsm
