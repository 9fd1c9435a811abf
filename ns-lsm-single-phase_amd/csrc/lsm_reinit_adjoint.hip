// Reverse mode (adjoint) of the Sussman reinitialisation
// (reinitialisation_test2.F90:42-210 with calc_NormGrad_d.F90:39-2446 and the
// mass correction of global_lsm.F90:120-133,200-219; the reference has no
// adjoint).
//
// Forward, per pseudo-time iteration and RK stage r, at every node with the
// iteration's update mask set (|grad d^n| - 1 > eps, frozen):
//     d_out = (c1 d^n + c2 d_in) - (c3 dtau) rhs,
//     rhs   = s(phi0) (|grad d_in| - 1) - lambda delta(d_in) |grad d_in|,
// and d_out = d_in elsewhere.  Where lambda is applied (delta^2 |grad d| >=
// eps_sims) it equals -s (1 - |grad d|) / (delta |grad d|), so rhs is
// identically zero there and so is its derivative; elsewhere lambda = 0 and
// rhs = s (|grad d| - 1).  |grad d| is linear in the stencil values once the
// ENO2 / biased branch, the upwind side chosen by phi0 and the boundary
// formulas are fixed, which the adjoint takes as in the forward; at
// |grad d| = 0 its derivative is taken as 0.  The iteration count, the update
// mask and every sign switch are likewise taken from the forward.
//
// Per stage (reverse order) two kernels.  `radj_partials` forms at each node
// the derivative of its update with respect to the d values of its stencil,
// scaled by the node's adjoint, in compact form: per direction one weight
// (adjoint x d|grad d|/d(D_dir d) x spacing) and a 4-bit code of the branch
// the direction's difference formula took (ENO2 result and second
// differences, or the ENO2Boundary case), 26 bytes per node instead of the
// 15 stencil coefficients; nodes whose update does not depend on the stencil
// (mask off, or lambda applied) write only a zero code.  `radj_gather` sums,
// at each node, the contributions of the nodes whose stencil holds it (slots:
// offsets -2..2 inside, 0..3 away from a face on a face), rebuilding each
// coefficient from the code, plus the linear parts.  The adjoint with respect
// to s(phi0) is accumulated per node inside the interface band (elsewhere
// s'(phi0) = 0) and turned into dJ/dphi0 at the end.  Deterministic (no
// atomics).
#include "lsm_internal.hpp"
#include "lsm_reinit_common.hpp"

namespace lsm {

namespace {

using rn::Geo;
using rn::make_geo;
using rn::node_of;

// |grad d| at node (i,j,k) (the formula of normgrad_node), the direction
// weights w_d = (grad d . grad xi_d / |grad d|) x spacing and the branch codes
// of the three difference formulas (4 bits each)
__device__ double normgrad_branch(const Geo &g, const double *__restrict__ d, const double *__restrict__ phi0,
                                  double phi0s, int order_bc, bool eno_bc, int i, int j, int k, long l,
                                  double w[3], int &code)
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
    const double gx = (g.m[0][l] * gr[0] + g.m[3][l] * gr[1]) + g.m[6][l] * gr[2];
    const double gy = (g.m[1][l] * gr[0] + g.m[4][l] * gr[1]) + g.m[7][l] * gr[2];
    const double gz = (g.m[2][l] * gr[0] + g.m[5][l] * gr[1]) + g.m[8][l] * gr[2];
    const double ng = sqrt((gx * gx + gy * gy) + gz * gz);
    const double inv = ng > 0.0 ? 1.0 / ng : 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        w[a] = ((gx * g.m[3 * a][l] + gy * g.m[3 * a + 1][l]) + gz * g.m[3 * a + 2][l]) * inv * scale[a];
    return ng;
}

constexpr uint16_t kLive = 0x8000;   // the node contributes to its stencil

struct RPartArgs {
    Geo g;
    const double *din, *phi0, *lam;
    const uint8_t *conv;
    const int32_t *iters;
    int it, order_bc, eno_bc;
    double c3dt, eps;
    double *q[3];      // direction weights times the node's scaled adjoint
    uint16_t *code;    // kLive | branch codes
    double *adj_s;     // accumulated dJ/ds(phi0) (inside the band)
    long n;
};

__global__ __launch_bounds__(256) void radj_partials(RPartArgs a)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.n) return;
    if (a.it >= *a.iters) return;   // the iteration did not run
    uint16_t code = 0;
    if (a.conv[l]) {
        int i, j, k;
        node_of(a.g, l, i, j, k);
        const Geo &g = a.g;
        const double p0s = a.phi0[(long)(g.ni - 2) + g.sj * (g.nj - 2) + g.sk * (g.nk - 2)];
        double w[3];
        int br;
        const double ng = normgrad_branch(g, a.din, a.phi0, p0s, a.order_bc, a.eno_bc != 0, i, j, k, l, w, br);
        const double p0 = a.phi0[l];
        const double s = sign_phi0(p0, a.eps);
        const double del = deltaf(a.din[l], a.eps);
        if (fabs((del * del) * ng) < kEpsSims) {   // lambda = 0: rhs = s (|grad d| - 1)
            const double mu = -a.c3dt * a.lam[l];
            if (fabs(p0) <= a.eps) a.adj_s[l] += mu * (ng - 1.0);
            const double ag = mu * s;
            if (ag != 0.0 && ng > 0.0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) a.q[d][l] = ag * w[d];
                code = (uint16_t)(kLive | br);
            }
        }
    }
    a.code[l] = code;
}

struct RGatherArgs {
    int ni, nj, nk;
    long sj, sk, n;
    const double *lam;
    const double *q[3];
    const uint16_t *code;
    const uint8_t *conv;
    const int32_t *iters;
    int it, stage, order_bc, eno_bc;
    double c1, c2;
    double *out;   // adjoint of the stage input
    double *adn;   // adjoint of the iteration input d^n, accumulated over stages 3, 2
};

__global__ __launch_bounds__(256) void radj_gather(RGatherArgs a)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const bool active = a.it < *a.iters;
    const double lam = a.lam[m];
    const bool upd = active && a.conv[m];
    double acc = upd ? a.c2 * lam : lam;
    if (active) {
        const int i = (int)(m % a.ni);
        const long r = m / a.ni;
        const int idx[3] = {i, (int)(r % a.nj), (int)(r / a.nj)};
        const int nn[3] = {a.ni, a.nj, a.nk};
        const long st[3] = {1, a.sj, a.sk};
        const bool j_on = idx[1] == 0 || idx[1] == a.nj - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
            for (int o = -3; o <= 3; ++o) {
                const int ia = idx[d] - o;   // index of the node whose offset-o entry is this node
                if (ia < 0 || ia > nn[d] - 1) continue;
                int slot, kind;              // kind: 0 inside, 1 low face, 2 high face
                if (ia == 0) { slot = o; kind = 1; }
                else if (ia == nn[d] - 1) { slot = -o; kind = 2; }
                else { slot = o + 2; kind = 0; }
                if (slot < 0 || slot > (kind == 0 ? 4 : 3)) continue;
                const long l = m - (long)o * st[d];
                const int c = a.code[l];
                if (!(c & kLive)) continue;
                const int br = (c >> (4 * d)) & 15;
                double e[5];
                if (kind == 0) {
                    eno2_coef(br, e);
                } else if (d == 0 && ia == 0 && idx[2] == 0 && !j_on) {
                    e[0] = -3.0; e[1] = 4.0; e[2] = -1.0; e[3] = 0.0;   // edge i = 1, k = 1
                } else {
                    biased_coef(br, a.order_bc, kind == 1 ? 1 : -1, a.eno_bc != 0, e);
                }
                acc += a.q[d][l] * e[slot];
            }
        }
    }
    const double lin = upd ? a.c1 * lam : 0.0;
    if (a.stage == 3) { a.adn[m] = lin; a.out[m] = acc; }
    else if (a.stage == 2) { a.adn[m] += lin; a.out[m] = acc; }
    else a.out[m] = acc + lin + a.adn[m];
}

// adj_phi0 = lam + adj_s * s'(phi0)
__global__ __launch_bounds__(256) void radj_final(long n, const double *lam, const double *adj_s,
                                                  const double *phi0, double eps, double *adj_phi0)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    adj_phi0[l] = lam[l] + adj_s[l] * sign_phi0_deriv(phi0[l], eps);
}

__global__ __launch_bounds__(256) void radj_zero(long n, double *p)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) p[l] = 0.0;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }
size_t fbytes(const Blk &b) { return ((size_t)b.n * sizeof(double) + 255) & ~(size_t)255; }

}  // namespace

size_t reinit_eno2_adjoint_work(const Blk &b, int itermax)
{
    const size_t code_bytes = ((size_t)b.n * sizeof(uint16_t) + 255) & ~(size_t)255;
    return reinit_trace_work(b, itermax) + 7 * fbytes(b) + code_bytes + 4096;
}

int launch_reinit_eno2_adjoint(const Blk &b, const LsmReinitParams &p, const double *phi0,
                               const double *adj_out, double *adj_phi0, double *out, Carver &w,
                               hipStream_t s, const ReinitTrace *pre)
{
    ReinitTrace own;
    int rc;
    if (!pre) {
        if ((rc = launch_reinit_eno2_trace(b, p, phi0, own, out, w, s))) return rc;
        pre = &own;
    }
    const ReinitTrace &t = *pre;
    const size_t fb = (size_t)b.n * sizeof(double);
    double *L0 = (double *)w.take(fb), *L1 = (double *)w.take(fb);
    double *adn = (double *)w.take(fb), *adj_s = (double *)w.take(fb);
    double *Q[3];
    for (auto &q : Q) q = (double *)w.take(fb);
    uint16_t *code = (uint16_t *)w.take((size_t)b.n * sizeof(uint16_t));
    if (!L0 || !L1 || !adn || !adj_s || !Q[0] || !Q[1] || !Q[2] || !code)
        return fail(LSM_ERR_WORKSPACE, "reinit adjoint workspace too small");
    if ((rc = launch_copy(adj_out, L0, b.n, s))) return rc;
    radj_zero<<<nblk(b.n), 256, 0, s>>>(b.n, adj_s);
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const Geo g = make_geo(b);
    double *lam = L0, *other = L1;
    for (int it = t.nit - 1; it >= 0; --it) {
        const double *sin_[3] = {t.D[it], t.A[it], t.B[it]};
        for (int st = 2; st >= 0; --st) {
            RPartArgs pa{g, sin_[st], phi0, lam, t.conv[it], t.iters, it, p.order_bc, p.eno_bc,
                         C[st][2] * p.rn_timestep, p.epslsm, {Q[0], Q[1], Q[2]}, code, adj_s, b.n};
            radj_partials<<<nblk(b.n), 256, 0, s>>>(pa);
            RGatherArgs ga{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, b.n, lam, {Q[0], Q[1], Q[2]}, code,
                           t.conv[it], t.iters, it, st + 1, p.order_bc, p.eno_bc, C[st][0], C[st][1], other, adn};
            radj_gather<<<nblk(b.n), 256, 0, s>>>(ga);
            double *tmp = lam; lam = other; other = tmp;
        }
        if ((rc = check_hip(hipGetLastError(), "reinit adjoint kernels"))) return rc;
    }
    radj_final<<<nblk(b.n), 256, 0, s>>>(b.n, lam, adj_s, phi0, p.epslsm, adj_phi0);
    return check_hip(hipGetLastError(), "radj_final");
}

}  // namespace lsm
