// Reverse mode (adjoint) of one SSP-RK3 step of the WENO3 level-set advection
// (ad_RKTVD3.F90:1-88 with calc_RH_AD_WENO3.F90:1-407 and bcond_lsm.F90:2-568).
//
// Given lam = dJ/dphi^(n+1), returns dJ/dphi^n and accumulates dJ/d(U/J).
// The three forward stages are recomputed first; then, for stage s = 3, 2, 1:
//   1. bcond_lsm transposed: its three face phases in reverse order (k, j, i),
//      each face value's adjoint handed to the two inner nodes it was
//      extrapolated from (BigPhi switch and btype as taken in the forward);
//   2. per node, the WENO3 right-hand side differentiated with respect to the
//      five stencil values of each direction (nonlinear weights included, the
//      upwind choice and the forced boundary stencils as in the forward), the
//      partials scaled by the node's adjoint and written out (15 fields), and
//      the derivative with respect to the node's U/J accumulated;
//   3. per node, a gather of the partials of the nodes whose stencil holds it,
//      plus the linear part c2 (and c1 into dJ/dphi^n).
// Everything is deterministic (no atomics).
#include <stdlib.h>

#include "lsm_internal.hpp"
#include "lsm_multi.hpp"

namespace lsm {

namespace {

// GetWENO3Reconstruction (AdvectionMethods.F90:195-245) and its gradient with
// respect to the stencil values f0..f3 on the upwind branch the forward takes
__device__ __forceinline__ double weno3_grad(double f0, double f1, double f2, double f3, double ws,
                                             double g[4])
{
    const bool pos = ws > 0.0;
    const double fL = pos ? f0 : f3;
    const double fC = pos ? f1 : f2;
    const double fR = pos ? f2 : f1;
    const double d0 = 2.0 / 3.0, d1 = 1.0 / 3.0, e = 0.000001;
    const double b0 = (fC - fR) * (fC - fR);
    const double b1 = (fL - fC) * (fL - fC);
    const double s0 = b0 + e, s1 = b1 + e;
    const double a0 = d0 / (s0 * s0);
    const double a1 = d1 / (s1 * s1);
    const double w0 = a0 / (a0 + a1);
    const double w1 = a1 / (a0 + a1);
    const double u0 = 0.5 * fC + 0.5 * fR;
    const double u1 = (-0.5) * fL + 1.5 * fC;
    // d w0 = w0 w1 (d ln a0 - d ln a1),  d ln a_k = -2 d b_k / s_k
    const double du = u0 - u1, ww = w0 * w1;
    const double q0 = 4.0 * (fC - fR) / s0, q1 = 4.0 * (fL - fC) / s1;
    const double gL = -0.5 * w1 + du * ww * q1;
    const double gC = 0.5 * w0 + 1.5 * w1 - du * ww * (q0 + q1);
    const double gR = 0.5 * w0 + du * ww * q0;
    if (pos) { g[0] = gL; g[1] = gC; g[2] = gR; g[3] = 0.0; }
    else { g[3] = gL; g[2] = gC; g[1] = gR; g[0] = 0.0; }
    return u0 * w0 + u1 * w1;
}

// one direction of calc_RH_AD_WENO3.F90:195-254: dL/dphi at offsets -2..2
// (L = ws d (Phi+ - Phi-)) and Phi+ - Phi-
__device__ __forceinline__ double weno3_dir_grad(double m2, double m1, double c0, double p1, double p2,
                                                 double ws, bool forced_minus, bool forced_plus,
                                                 double d, double gl[5])
{
    double gp[4], gm[4];
    const double fp = weno3_grad(m1, c0, p1, forced_plus ? 0.0 : p2, forced_plus ? 1.0 : ws, gp);
    const double fm = weno3_grad(forced_minus ? 0.0 : m2, m1, c0, p1, forced_minus ? -1.0 : ws, gm);
    const double k = ws * d;
    gl[0] = k * (-gm[0]);
    gl[1] = k * (gp[0] - gm[1]);
    gl[2] = k * (gp[1] - gm[2]);
    gl[3] = k * (gp[2] - gm[3]);
    gl[4] = k * gp[3];
    return fp - fm;
}

struct AdjGeo {
    int ni, nj, nk;
    long sj, sk, n;
};

__device__ __forceinline__ void node3(const AdjGeo &g, long l, int &i, int &j, int &k)
{
    i = (int)(l % g.ni);
    const long r = l / g.ni;
    j = (int)(r % g.nj);
    k = (int)(r / g.nj);
}

// (2) partials of the stage right-hand side, scaled by the node's adjoint
struct PartArgs {
    AdjGeo g;
    const double *prev, *lam, *U1, *U2, *U3, *aj;
    const int32_t *mask;
    double *P;                    // [15][n]: direction d, offset o -> P[(5 d + o + 2) n + l]
    double *aU1, *aU2, *aU3;      // accumulated dJ/dU (may be NULL)
    double c3, dt, dd[3];
};

__global__ __launch_bounds__(256) void adj_partials(PartArgs a)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.g.n) return;
    int i, j, k;
    node3(a.g, l, i, j, k);
    const int idx[3] = {i, j, k};
    const int nn[3] = {a.g.ni, a.g.nj, a.g.nk};
    const long st[3] = {1, a.g.sj, a.g.sk};
    const bool interior = i >= 1 && i <= a.g.ni - 2 && j >= 1 && j <= a.g.nj - 2 && k >= 1 &&
                          k <= a.g.nk - 2;
    const bool upd = interior && (a.mask == nullptr || a.mask[l] > 0);
    const long n = a.g.n;
    if (!upd) {
        for (int q = 0; q < 15; ++q) a.P[q * n + l] = 0.0;
        return;
    }
    // phi~ = c1 phi^n + c2 phi - ((c3 aj) dt) rh  (ad_RKTVD3.F90:70-72)
    const double m = a.lam[l] * (-((a.c3 * a.aj[l]) * a.dt));
    const double U[3] = {a.U1[l], a.U2[l], a.U3[l]};
    double *aU[3] = {a.aU1, a.aU2, a.aU3};
    const double *p = a.prev;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long s = st[d];
        const bool fm = idx[d] == 1, fp = idx[d] == nn[d] - 2;
        const double m2 = fm ? 0.0 : p[l - 2 * s], p2 = fp ? 0.0 : p[l + 2 * s];
        double gl[5];
        const double dphi = weno3_dir_grad(m2, p[l - s], p[l], p[l + s], p2, U[d], fm, fp, a.dd[d], gl);
#pragma unroll
        for (int o = 0; o < 5; ++o) a.P[(5 * d + o) * n + l] = m * gl[o];
        if (aU[d]) aU[d][l] += m * (a.dd[d] * dphi);
    }
}

// (3) gather: adjoint of the stage input (and the c1 part of dJ/dphi^n)
struct GatherArgs {
    AdjGeo g;
    const double *lam, *P;
    const int32_t *mask;
    double *prev_out;   // adjoint of the stage input; NULL: add it to adj_n instead
    double *adj_n;      // accumulated dJ/dphi^n
    double c1, c2;
};

__global__ __launch_bounds__(256) void adj_gather(GatherArgs a)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.g.n) return;
    int i, j, k;
    node3(a.g, l, i, j, k);
    const int idx[3] = {i, j, k};
    const int nn[3] = {a.g.ni, a.g.nj, a.g.nk};
    const long st[3] = {1, a.g.sj, a.g.sk};
    const long n = a.g.n;
    const double lam = a.lam[l];
    const bool upd = a.mask == nullptr || a.mask[l] > 0;
    double acc = upd ? a.c2 * lam : lam;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int o = -2; o <= 2; ++o) {
            const int src = idx[d] - o;   // node whose offset-o stencil entry is this node
            if (src >= 1 && src <= nn[d] - 2) acc += a.P[(5 * d + o + 2) * n + l - (long)o * st[d]];
        }
    }
    const double lin = upd ? a.c1 * lam : 0.0;
    if (a.prev_out) {
        a.prev_out[l] = acc;
        a.adj_n[l] += lin;
    } else {
        a.adj_n[l] += lin + acc;
    }
}

// (1) bcond_lsm transposed, one face phase (axis 2, 1, 0 in that order); the
// BigPhi switch is evaluated on the stage's final values, which at the two
// inner nodes of a face equal the values the face saw in the forward
struct BcAdjArgs {
    AdjGeo g;
    const double *phi;   // forward stage result after bcond_lsm
    double *lam;         // adjoint, updated in place
    int bt[6];
    bool flo[3], fhi[3];
    double big;
};

__device__ __forceinline__ void bc_coefs(int bt, double p2, double big, bool is_im, double &a, double &c)
{
    if (bt == 0) { a = 1.0; c = 0.0; return; }
    if (bt == 1) {
        const double a2 = 4.0 / 3.0, b2 = -1.0 / 3.0, a1 = 1.0, b1 = 0.0;
        const double ex = (1.0 + sign1(fabs(p2) - big)) / 2.0;
        a = a2 + ex * (a1 - a2);
        c = b2 + ex * (b1 - b2);
        return;
    }
    if (is_im) { a = 2.0; c = -1.0; return; }
    a = 1.0; c = 0.0;   // not reached: type 2 leaves other faces alone
}

__device__ __forceinline__ void adj_bcond_phase_run(const BcAdjArgs &a, int axis, int bid);

struct BcAdjLaunch {
    BcAdjArgs a;
    int axis;
};

struct KBcAdj {
    using Args = BcAdjLaunch;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &x, int bid, int) { adj_bcond_phase_run(x.a, x.axis, bid); }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &x);
};

__global__ __launch_bounds__(256) void adj_bcond_phase(BcAdjArgs a, int axis)
{
    adj_bcond_phase_run(a, axis, (int)blockIdx.x);
}

void KBcAdj::one(unsigned g, dim3 b, hipStream_t s, const Args &x) { adj_bcond_phase<<<g, b, 0, s>>>(x.a, x.axis); }

__device__ __forceinline__ void adj_bcond_phase_run(const BcAdjArgs &a, int axis, int bid)
{
    const long t = (long)bid * 256 + threadIdx.x;
    const AdjGeo &g = a.g;
    int n1, n2, nax;
    long s, s1, s2;
    if (axis == 0) { n1 = g.nj; n2 = g.nk; nax = g.ni; s = 1; s1 = g.sj; s2 = g.sk; }
    else if (axis == 1) { n1 = g.ni; n2 = g.nk; nax = g.nj; s = g.sj; s1 = 1; s2 = g.sk; }
    else { n1 = g.ni; n2 = g.nj; nax = g.nk; s = g.sk; s1 = 1; s2 = g.sj; }
    if (t >= (long)n1 * n2) return;
    const long base = (t % n1) * s1 + (t / n1) * s2;
    // faces of one phase do not interact (extents >= 4); the high face first, as
    // the reverse of the forward order
    const int bt_lo = a.bt[2 * axis], bt_hi = a.bt[2 * axis + 1];
    if (a.fhi[axis] && !(bt_hi == 2 && axis != 0)) {
        const long l = base + (long)(nax - 1) * s;
        double ca, cc;
        bc_coefs(bt_hi, a.phi[l - 2 * s], a.big, axis == 0, ca, cc);
        const double v = a.lam[l];
        a.lam[l - s] += ca * v;
        a.lam[l - 2 * s] += cc * v;
        a.lam[l] = 0.0;
    }
    if (a.flo[axis] && bt_lo != 2) {
        const long l = base;
        double ca, cc;
        bc_coefs(bt_lo, a.phi[l + 2 * s], a.big, false, ca, cc);
        const double v = a.lam[l];
        a.lam[l + s] += ca * v;
        a.lam[l + 2 * s] += cc * v;
        a.lam[l] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Fused form of steps (2) and (3): the partials stay on chip.  A workgroup of
// 64 (i) x R (j) threads owns the 60 x (R - 4) nodes in its middle and marches
// up a chunk of k planes; the two outer lanes/rows on each side only compute
// partials for their neighbours.  Per plane: the stage input with a two-node
// halo is staged in LDS, every thread forms the partials of its node, the i and
// j partials go to LDS and are gathered by the nodes they point at, and the k
// partials go to a five-plane register window of the thread's own column, so a
// node's adjoint is complete two planes later.  The WENO3 gradient uses the
// one-quotient form of the reconstruction (one reciprocal per face value).
// ---------------------------------------------------------------------------

// gradient of Phi = ((fC + fR) A1 + (1.5 fC - 0.5 fL) A0) / (2 A1 + A0),
// A_k = ((diff_k)^2 + 1e-6)^2 (the reconstruction of weno3_fast) with respect to
// f0..f3 on the forward's upwind branch
__device__ __forceinline__ double weno3_grad_fast(double f0, double f1, double f2, double f3, double ws,
                                                  double g[4])
{
#pragma clang fp contract(fast)
    const bool pos = ws > 0.0;
    const double fL = pos ? f0 : f3;
    const double fC = pos ? f1 : f2;
    const double fR = pos ? f2 : f1;
    const double e = 0.000001;
    const double t0 = fC - fR, t1 = fL - fC;
    const double s0 = t0 * t0 + e, s1 = t1 * t1 + e;
    const double A0 = s0 * s0, A1 = s1 * s1;
    const double u1 = 1.5 * fC - 0.5 * fL;
    const double N = (fC + fR) * A1 + u1 * A0;
    const double D = 2.0 * A1 + A0;
    double r = __builtin_amdgcn_rcp(D);
    double er = fma(-D, r, 1.0);
    r = fma(r, er, r);
    er = fma(-D, r, 1.0);
    r = fma(r, er, r);
    const double phi = N * r;
    // dA0 = 4 s0 t0 (dfC - dfR), dA1 = 4 s1 t1 (dfL - dfC)
    const double k0 = 4.0 * s0 * t0, k1 = 4.0 * s1 * t1;
    const double cA0 = u1 - phi, cA1 = (fC + fR) - 2.0 * phi;   // dPhi = (dN - phi dD) / D
    const double gL = (cA1 * k1 - 0.5 * A0) * r;
    const double gC = (A1 + 1.5 * A0 + cA0 * k0 - cA1 * k1) * r;
    const double gR = (A1 - cA0 * k0) * r;
    if (pos) { g[0] = gL; g[1] = gC; g[2] = gR; g[3] = 0.0; }
    else { g[3] = gL; g[2] = gC; g[1] = gR; g[0] = 0.0; }
    return phi;
}

__device__ __forceinline__ double weno3_dir_grad_fast(double m2, double m1, double c0, double p1, double p2,
                                                      double ws, bool forced_minus, bool forced_plus,
                                                      double d, double gl[5])
{
    double gp[4], gm[4];
    const double fp = weno3_grad_fast(m1, c0, p1, forced_plus ? 0.0 : p2, forced_plus ? 1.0 : ws, gp);
    const double fm = weno3_grad_fast(forced_minus ? 0.0 : m2, m1, c0, p1, forced_minus ? -1.0 : ws, gm);
    const double k = ws * d;
    gl[0] = k * (-gm[0]);
    gl[1] = k * (gp[0] - gm[1]);
    gl[2] = k * (gp[1] - gm[2]);
    gl[3] = k * (gp[2] - gm[3]);
    gl[4] = k * gp[3];
    return fp - fm;
}

constexpr int AX = 64, AOX = 60;   // lanes per row, output nodes per row

struct FusedArgs {
    AdjGeo g;
    const double *prev, *lam, *U1, *U2, *U3, *aj;
    const int32_t *mask;
    double *prev_out;   // adjoint of the stage input; NULL: add it to adj_n instead
    double *adj_n;
    double *aU1, *aU2, *aU3;
    double c1, c2, c3, dt, dd[3];
    int tiles_i, tiles_j, kc;
};

template <int R>
__global__ __launch_bounds__(AX * R) void adj_stage_fused(FusedArgs a)
{
    constexpr int OY = R - 4;
    const AdjGeo &g = a.g;
    const int b = blockIdx.x;
    const int ti = b % a.tiles_i, t2 = b / a.tiles_i;
    const int tj = t2 % a.tiles_j, tk = t2 / a.tiles_j;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * AX + tx;
    const int i = ti * AOX - 2 + tx, j = tj * OY - 2 + ty;
    const int kb = tk * a.kc, ke = min(kb + a.kc - 1, g.nk - 1);
    const bool inb = i >= 0 && i < g.ni && j >= 0 && j < g.nj;
    const bool outq = inb && tx >= 2 && tx < 2 + AOX && ty >= 2 && ty < 2 + OY;
    const bool inter_ij = i >= 1 && i <= g.ni - 2 && j >= 1 && j <= g.nj - 2;
    const int ic = min(max(i, 0), g.ni - 1), jc = min(max(j, 0), g.nj - 1);
    const long col = (long)ic + g.sj * jc;
    auto kcl = [&](int kk) -> long { return g.sk * (long)min(max(kk, 0), g.nk - 1); };

    __shared__ double sp[R + 4][AX + 4];      // stage input of the plane, two-node halo
    __shared__ double Pj[5][R][AX];           // j partials (offset -2..2) of each node; the i
                                              // partials stay in registers and move by lane shuffles

    const double *__restrict__ prev = a.prev;
    double q0 = prev[col + kcl(kb - 4)], q1 = prev[col + kcl(kb - 3)], q2 = prev[col + kcl(kb - 2)];
    double q3 = prev[col + kcl(kb - 1)], q4 = prev[col + kcl(kb)];
    double kacc0 = 0.0, kacc1 = 0.0, kacc2 = 0.0, kacc3 = 0.0, kacc4 = 0.0;   // planes p-2 .. p+2
    double pend0 = 0.0, pend1 = 0.0;                                          // planes p-2, p-1
    const int ncell = (R + 4) * (AX + 4);
    for (int p = kb - 2; p <= ke + 2; ++p) {
        // q0..q4 = prev(p-2 .. p+2) of this column
        const bool live = p >= 0 && p < g.nk;
        for (int c = tid; c < ncell; c += AX * R) {
            const int rr = c / (AX + 4), cc = c % (AX + 4);
            const int ii = min(max(ti * AOX - 4 + cc, 0), g.ni - 1);
            const int jj = min(max(tj * OY - 4 + rr, 0), g.nj - 1);
            sp[rr][cc] = prev[(long)ii + g.sj * jj + kcl(p)];
        }
        double lam = 0.0;
        bool upd = false;
        double U[3] = {0.0, 0.0, 0.0}, ajv = 0.0;
        if (live && inb) {
            const long l = col + g.sk * (long)p;
            lam = a.lam[l];
            upd = a.mask == nullptr || a.mask[l] > 0;
            U[0] = a.U1[l]; U[1] = a.U2[l]; U[2] = a.U3[l];
            ajv = a.aj[l];
        }
        __syncthreads();
        const bool part = live && inb && inter_ij && p >= 1 && p <= g.nk - 2 && upd;
        double pi0 = 0.0, pi1 = 0.0, pi2 = 0.0, pi3 = 0.0, pi4 = 0.0;
        if (part) {
            // one direction at a time, each written out before the next is formed
            const double m = lam * (-((a.c3 * ajv) * a.dt));
            const bool own = outq && p >= kb && p <= ke && a.aU1 != nullptr;
            const long l = col + g.sk * (long)p;
            const int cx = tx + 2, cy = ty + 2;
            double gl[5];
            double dph = weno3_dir_grad_fast(sp[cy][cx - 2], sp[cy][cx - 1], sp[cy][cx], sp[cy][cx + 1],
                                             sp[cy][cx + 2], U[0], i == 1, i == g.ni - 2, a.dd[0], gl);
            pi0 = m * gl[0]; pi1 = m * gl[1]; pi2 = m * gl[2]; pi3 = m * gl[3]; pi4 = m * gl[4];
            if (own) a.aU1[l] += m * (a.dd[0] * dph);
            __builtin_amdgcn_sched_barrier(0);
            dph = weno3_dir_grad_fast(sp[cy - 2][cx], sp[cy - 1][cx], sp[cy][cx], sp[cy + 1][cx],
                                      sp[cy + 2][cx], U[1], j == 1, j == g.nj - 2, a.dd[1], gl);
#pragma unroll
            for (int o = 0; o < 5; ++o) Pj[o][ty][tx] = m * gl[o];
            if (own) a.aU2[l] += m * (a.dd[1] * dph);
            __builtin_amdgcn_sched_barrier(0);
            dph = weno3_dir_grad_fast(q0, q1, q2, q3, q4, U[2], p == 1, p == g.nk - 2, a.dd[2], gl);
            kacc0 += m * gl[0]; kacc1 += m * gl[1]; kacc2 += m * gl[2]; kacc3 += m * gl[3];
            kacc4 += m * gl[4];
            if (own) a.aU3[l] += m * (a.dd[2] * dph);
        } else {
#pragma unroll
            for (int o = 0; o < 5; ++o) Pj[o][ty][tx] = 0.0;
        }
        // i partial o of lane tx - (o - 2) (a wavefront is one tile row)
        const double gi[5] = {__shfl(pi0, (tx + 2) & (AX - 1)), __shfl(pi1, (tx + 1) & (AX - 1)), pi2,
                              __shfl(pi3, (tx - 1) & (AX - 1)), __shfl(pi4, (tx - 2) & (AX - 1))};
        __syncthreads();
        double pend2 = 0.0;
        if (outq && live && p >= kb && p <= ke) {
            // partial of node tx - (o - 2) with respect to this node
            double sacc = upd ? a.c2 * lam : lam;
#pragma unroll
            for (int o = 0; o < 5; ++o) sacc += gi[o] + Pj[o][ty - (o - 2)][tx];
            pend2 = sacc;
            if (upd && a.prev_out) a.adj_n[col + g.sk * (long)p] += a.c1 * lam;
            else if (upd) pend2 += a.c1 * lam;
        }
        // plane p - 2 has all its k partials now
        const int pf = p - 2;
        if (outq && pf >= kb && pf <= ke) {
            const double tot = pend0 + kacc0;
            const long l = col + g.sk * (long)pf;
            if (a.prev_out) a.prev_out[l] = tot;
            else a.adj_n[l] += tot;
        }
        kacc0 = kacc1; kacc1 = kacc2; kacc2 = kacc3; kacc3 = kacc4; kacc4 = 0.0;
        pend0 = pend1; pend1 = pend2;
        q0 = q1; q1 = q2; q2 = q3; q3 = q4; q4 = prev[col + kcl(p + 3)];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The same stage adjoint with each thread doing only the work its node needs
// and the loads of the next plane in flight while a plane computes.  Tiles and
// k march as adj_stage_fused; rows 0, 1, R-2, R-1 of the tile (whole
// wavefronts) only supply the j partials of their nodes, lanes 0, 1, 62, 63
// only the i partials, and the two planes below and above the k chunk only
// the k partials of the tile's own columns.  The i partials of a row are
// summed across lanes at once (shuffles), the j partials pass through LDS,
// the k partials stay in the column's register window.
// ---------------------------------------------------------------------------
template <int R>
struct KAdjLean {
    using Args = FusedArgs;
    static constexpr int kThreads = AX * R;
    __device__ __forceinline__ static void run(const Args &a, int bid, int nb);
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

template <int R>
__global__ __launch_bounds__(AX * R) void adj_stage_lean(FusedArgs a)
{
    KAdjLean<R>::run(a, (int)blockIdx.x, (int)gridDim.x);
}

template <int R>
void KAdjLean<R>::one(unsigned g, dim3 b, hipStream_t s, const Args &a)
{
    adj_stage_lean<R><<<g, b, 0, s>>>(a);
}

template <int R>
__device__ __forceinline__ void KAdjLean<R>::run(const FusedArgs &a, int bid, int nb)
{
    constexpr int OY = R - 4;
    constexpr int NCELL = (R + 4) * (AX + 4);
    const AdjGeo &g = a.g;
    const int ntile = a.tiles_i * a.tiles_j * ((g.nk + a.kc - 1) / a.kc);
    // consecutive tiles on one XCD (workgroup b runs on XCD b % 8)
    const int b = (nb & 7) == 0 ? (int)((bid & 7) * (nb >> 3) + (bid >> 3)) : bid;
    if (b >= ntile) return;
    const int ti = b % a.tiles_i, t2 = b / a.tiles_i;
    const int tj = t2 % a.tiles_j, tk = t2 / a.tiles_j;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * AX + tx;
    const int i = ti * AOX - 2 + tx, j = tj * OY - 2 + ty;
    const int kb = tk * a.kc, ke = min(kb + a.kc - 1, g.nk - 1);
    const bool hrow = ty < 2 || ty >= R - 2;               // uniform per wavefront
    const bool hlane = tx < 2 || tx >= 2 + AOX;
    const bool inb = i >= 0 && i < g.ni && j >= 0 && j < g.nj;
    const bool own = inb && !hrow && !hlane;
    const bool inter_ij = i >= 1 && i <= g.ni - 2 && j >= 1 && j <= g.nj - 2;
    const int ic = min(max(i, 0), g.ni - 1), jc = min(max(j, 0), g.nj - 1);
    const long col = (long)ic + g.sj * jc;
    auto kcl = [&](int kk) -> long { return g.sk * (long)min(max(kk, 0), g.nk - 1); };
    // the two LDS cells of the stage-input plane this thread loads
    const int c0 = tid, c1 = tid + AX * R;
    auto cell = [&](int c, int p) -> long {
        const int rr = c / (AX + 4), cc = c - rr * (AX + 4);
        const int ii = min(max(ti * AOX - 4 + cc, 0), g.ni - 1);
        const int jj = min(max(tj * OY - 4 + rr, 0), g.nj - 1);
        return (long)ii + g.sj * jj + kcl(p);
    };

    __shared__ double sp[2][R + 4][AX + 4];   // stage input of the plane, two-node halo
    __shared__ double Pj[5][R][AX];           // j partials (offset -2..2) of each node

    const double *__restrict__ prev = a.prev;
    const double *__restrict__ lamp = a.lam;
    // operands of plane p, loaded one plane ahead
    auto need = [&](int p, bool &ni_, bool &nj_, bool &nk_) {
        const bool core = p >= kb && p <= ke, live = p >= 0 && p < g.nk;
        ni_ = live && core && inb && !hrow;
        nj_ = live && core && inb && !hlane;
        nk_ = live && own;
    };
    // k window of the column: q0..q4 = prev(p-2 .. p+2) (clamped), q4 loaded
    // one plane ahead
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0;
    if (own) {
        q0 = prev[col + kcl(kb - 4)]; q1 = prev[col + kcl(kb - 3)]; q2 = prev[col + kcl(kb - 2)];
        q3 = prev[col + kcl(kb - 1)];
    }
    double nlam = 0.0, naj = 0.0, nu1 = 0.0, nu2 = 0.0, nu3 = 0.0, nq = 0.0, ns0 = 0.0, ns1 = 0.0;
    int nupd = 1;
    auto fetch = [&](int p) {
        bool fi, fj, fk;
        need(p, fi, fj, fk);
        const long l = col + g.sk * (long)min(max(p, 0), g.nk - 1);
        if (fi || fj || fk) {
            nlam = lamp[l];
            naj = a.aj[l];
            nupd = a.mask == nullptr || a.mask[l] > 0;
        }
        if (fi) nu1 = a.U1[l];
        if (fj) nu2 = a.U2[l];
        if (fk) nu3 = a.U3[l];
        if (own) nq = prev[col + kcl(p + 2)];
        if (p >= kb && p <= ke) {
            ns0 = prev[cell(c0, p)];
            if (c1 < NCELL) ns1 = prev[cell(c1, p)];
        }
    };
    fetch(kb - 2);
    double kacc0 = 0.0, kacc1 = 0.0, kacc2 = 0.0, kacc3 = 0.0, kacc4 = 0.0;   // planes p-2 .. p+2
    double pend0 = 0.0, pend1 = 0.0;                                          // planes p-2, p-1
    int buf = 0;
    for (int p = kb - 2; p <= ke + 2; ++p) {
        const bool core = p >= kb && p <= ke, live = p >= 0 && p < g.nk;
        bool fi, fj, fk;
        need(p, fi, fj, fk);
        const double lam = nlam, ajv = naj, u1 = nu1, u2 = nu2, u3 = nu3;
        const bool upd = nupd != 0;
        q4 = nq;
        if (core) {
            sp[buf][c0 / (AX + 4)][c0 % (AX + 4)] = ns0;
            if (c1 < NCELL) sp[buf][c1 / (AX + 4)][c1 % (AX + 4)] = ns1;
        }
        if (p < ke + 2) fetch(p + 1);
        const long l = col + g.sk * (long)p;
        const bool wr_u = own && core && a.aU1 != nullptr;
        double o1 = 0.0, o2 = 0.0, o3 = 0.0;
        if (wr_u) { o1 = a.aU1[l]; o2 = a.aU2[l]; o3 = a.aU3[l]; }
        const bool part = live && inb && inter_ij && p >= 1 && p <= g.nk - 2 && upd;
        const double m = lam * (-((a.c3 * ajv) * a.dt));
        double gsum = 0.0;
        double d1 = 0.0, d2 = 0.0, d3 = 0.0;
        if (core) {
            __syncthreads();
            const int cx = tx + 2, cy = ty + 2;
            const double (*P)[AX + 4] = sp[buf];
            if (!hrow) {
                double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0;
                if (part) {
                    double gl[5];
                    d1 = weno3_dir_grad_fast(P[cy][cx - 2], P[cy][cx - 1], P[cy][cx], P[cy][cx + 1], P[cy][cx + 2],
                                             u1, i == 1, i == g.ni - 2, a.dd[0], gl);
                    p0 = m * gl[0]; p1 = m * gl[1]; p2 = m * gl[2]; p3 = m * gl[3]; p4 = m * gl[4];
                }
                // i partial o of lane tx - (o - 2) with respect to this lane's node
                gsum = (((p2 + __shfl(p0, (tx + 2) & (AX - 1))) + __shfl(p1, (tx + 1) & (AX - 1))) +
                        __shfl(p3, (tx - 1) & (AX - 1))) + __shfl(p4, (tx - 2) & (AX - 1));
            }
            if (!hlane) {
                double gl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                if (part) {
                    d2 = weno3_dir_grad_fast(P[cy - 2][cx], P[cy - 1][cx], P[cy][cx], P[cy + 1][cx], P[cy + 2][cx],
                                             u2, j == 1, j == g.nj - 2, a.dd[1], gl);
                }
#pragma unroll
                for (int o = 0; o < 5; ++o) Pj[o][ty][tx] = part ? m * gl[o] : 0.0;
            }
        }
        if (own && part) {
            double gl[5];
            d3 = weno3_dir_grad_fast(q0, q1, q2, q3, q4, u3, p == 1, p == g.nk - 2, a.dd[2], gl);
            kacc0 += m * gl[0]; kacc1 += m * gl[1]; kacc2 += m * gl[2]; kacc3 += m * gl[3];
            kacc4 += m * gl[4];
        }
        double pend2 = 0.0;
        if (core) {
            __syncthreads();
            if (own && live) {
                double sacc = (upd ? a.c2 * lam : lam) + gsum;
#pragma unroll
                for (int o = 0; o < 5; ++o) sacc += Pj[o][ty - (o - 2)][tx];
                pend2 = sacc;
                if (upd && a.prev_out) a.adj_n[l] += a.c1 * lam;
                else if (upd) pend2 += a.c1 * lam;
            }
            if (wr_u && part) {
                a.aU1[l] = o1 + m * (a.dd[0] * d1);
                a.aU2[l] = o2 + m * (a.dd[1] * d2);
                a.aU3[l] = o3 + m * (a.dd[2] * d3);
            }
        }
        // plane p - 2 has all its k partials now
        const int pf = p - 2;
        if (own && pf >= kb && pf <= ke) {
            const double tot = pend0 + kacc0;
            const long lf = col + g.sk * (long)pf;
            if (a.prev_out) a.prev_out[lf] = tot;
            else a.adj_n[lf] += tot;
        }
        kacc0 = kacc1; kacc1 = kacc2; kacc2 = kacc3; kacc3 = kacc4; kacc4 = 0.0;
        pend0 = pend1; pend1 = pend2;
        q0 = q1; q1 = q2; q2 = q3; q3 = q4;
        buf ^= 1;
    }
}

struct FillArgs {
    double *p;
    long n;
    double v;
};

struct KFill {
    using Args = FillArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long l = (long)bid * 256 + threadIdx.x;
        if (l < a.n) a.p[l] = a.v;
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void adj_fill(double *p, long n, double v) { KFill::run(FillArgs{p, n, v}, (int)blockIdx.x, (int)gridDim.x); }

void KFill::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { adj_fill<<<g, b, 0, s>>>(a.p, a.n, a.v); }

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

size_t fbytes(const Blk &b) { return ((size_t)b.n * sizeof(double) + 255) & ~(size_t)255; }

}  // namespace

size_t rk3_adjoint_work(const Blk &b) { return 20 * fbytes(b) + 4096; }

namespace {

BcAdjArgs bc_adj_args(const Blk &b, const AdjGeo &g)
{
    BcAdjArgs ba{};
    ba.g = g;
    for (int q = 0; q < 6; ++q) ba.bt[q] = b.btype[q];
    ba.flo[0] = b.e.faces & LSM_FACE_I1; ba.fhi[0] = b.e.faces & LSM_FACE_IM;
    ba.flo[1] = b.e.faces & LSM_FACE_J1; ba.fhi[1] = b.e.faces & LSM_FACE_JM;
    ba.flo[2] = b.e.faces & LSM_FACE_K1; ba.fhi[2] = b.e.faces & LSM_FACE_KM;
    ba.big = b.big_phi;
    return ba;
}

}  // namespace

int launch_bcond_adjoint(const Blk &b, const double *phi, double *adj, hipStream_t s)
{
    const AdjGeo g{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, b.n};
    BcAdjArgs ba = bc_adj_args(b, g);
    ba.phi = phi;
    ba.lam = adj;
    const long lines[3] = {(long)b.e.nj * b.e.nk, (long)b.e.ni * b.e.nk, (long)b.e.ni * b.e.nj};
    for (int ax = 2; ax >= 0; --ax) {
        const int rc = klaunch<KBcAdj>(nblk(lines[ax]), dim3(256), s, BcAdjLaunch{ba, ax}, "adj_bcond_phase");
        if (rc) return rc;
    }
    return LSM_OK;
}

int launch_rk3_adjoint(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                       const double *U2, const double *U3, double dt, int mode, const double *adj_np1,
                       double *adj_n, double *adjU1, double *adjU2, double *adjU3, Carver &w,
                       hipStream_t s, const double *const *stages, bool adj_in_place)
{
    const size_t fb = (size_t)b.n * sizeof(double);
    const double *F[3];
    double *L[2];
    double *Fw[3] = {nullptr, nullptr, nullptr};
    if (!stages)
        for (auto &p : Fw) p = (double *)w.take(fb);
    // adj_in_place: the caller's adj_np1 is scratch the reverse sweep may overwrite
    L[0] = adj_in_place ? const_cast<double *>(adj_np1) : (double *)w.take(fb);
    L[1] = (double *)w.take(fb);
    for (int q = 0; q < 3; ++q) F[q] = stages ? stages[q] : Fw[q];
    if (!F[0] || !F[1] || !F[2] || !L[0] || !L[1]) return fail(LSM_ERR_WORKSPACE, "adjoint workspace too small");
    double *P = (double *)w.take(15 * fb);   // only the two-kernel form uses it
    int rc;
    // forward stages phi^(1), phi^(2), phi^(3)
    if (!stages &&
        (rc = launch_rk3_stages(b, phi_n, mask, U1, U2, U3, dt, mode, Fw[0], Fw[1], Fw[2], s))) return rc;
    if (!adj_in_place && (rc = launch_copy(adj_np1, L[0], b.n, s))) return rc;
    if ((rc = klaunch<KFill>(nblk(b.n), dim3(256), s, FillArgs{adj_n, b.n, 0.0}, "adj_fill"))) return rc;

    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const AdjGeo g{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, b.n};
    BcAdjArgs ba = bc_adj_args(b, g);
    const long lines[3] = {(long)b.e.nj * b.e.nk, (long)b.e.ni * b.e.nk, (long)b.e.ni * b.e.nj};

    // fused partials + gather (default) or the two-kernel form with the 15
    // partial fields in HBM
    // 2: the lean stage kernel (default); 16 / 8: the first fused form with
    // 16 / 8 rows; 0: the two-kernel form
    int fused = 2;
    if (const char *e = getenv("LSM_ADJ_FUSED")) fused = atoi(e);
    int kc = 16;
    if (const char *e = getenv("LSM_ADJ_KCHUNK")) { const int v = atoi(e); if (v > 0) kc = v; }
    if (fused != 2 && recording())
        return fail(LSM_ERR_UNSUPPORTED, "only the default advection adjoint kernel can be recorded");

    double *lam = L[0], *other = L[1];
    for (int st = 2; st >= 0; --st) {
        const double *phi_s = F[st];
        const double *prev = st == 0 ? phi_n : F[st - 1];
        // (1) bcond_lsm transposed
        ba.phi = phi_s;
        ba.lam = lam;
        for (int ax = 2; ax >= 0; --ax)
            if ((rc = klaunch<KBcAdj>(nblk(lines[ax]), dim3(256), s, BcAdjLaunch{ba, ax}, "adj_bcond_phase")))
                return rc;
        if (fused == 2) {
            FusedArgs fa{g, prev, lam, U1, U2, U3, b.aj, mask, st == 0 ? nullptr : other, adj_n,
                         adjU1, adjU2, adjU3, C[st][0], C[st][1], C[st][2], dt, {b.dc, b.de, b.dz},
                         (b.e.ni + AOX - 1) / AOX, (b.e.nj + 11) / 12, kc};
            const long nt = (long)fa.tiles_i * fa.tiles_j * ((b.e.nk + kc - 1) / kc);
            const unsigned grid = (unsigned)((nt + 7) / 8 * 8);
            KTimer kt(LSM_KT_ADJOINT_STAGE, s);
            if ((rc = klaunch<KAdjLean<16>>(grid, dim3(AX, 16, 1), s, fa, "adj_stage_lean"))) return rc;
        } else if (fused == 8 || fused == 16) {
            FusedArgs fa{g, prev, lam, U1, U2, U3, b.aj, mask, st == 0 ? nullptr : other, adj_n,
                         adjU1, adjU2, adjU3, C[st][0], C[st][1], C[st][2], dt, {b.dc, b.de, b.dz},
                         (b.e.ni + AOX - 1) / AOX, 0, kc};
            const int oy = fused - 4;
            fa.tiles_j = (b.e.nj + oy - 1) / oy;
            const unsigned grid = (unsigned)(fa.tiles_i * fa.tiles_j * ((b.e.nk + kc - 1) / kc));
            KTimer kt(LSM_KT_ADJOINT_STAGE, s);
            if (fused == 16) adj_stage_fused<16><<<grid, dim3(AX, 16, 1), 0, s>>>(fa);
            else adj_stage_fused<8><<<grid, dim3(AX, 8, 1), 0, s>>>(fa);
        } else {
            if (!P) return fail(LSM_ERR_WORKSPACE, "adjoint workspace too small");
            // (2) partials of the stage right-hand side
            PartArgs pa{g, prev, lam, U1, U2, U3, b.aj, mask, P, adjU1, adjU2, adjU3, C[st][2], dt,
                        {b.dc, b.de, b.dz}};
            adj_partials<<<nblk(b.n), 256, 0, s>>>(pa);
            // (3) gather into the adjoint of the stage input
            GatherArgs ga{g, lam, P, mask, st == 0 ? nullptr : other, adj_n, C[st][0], C[st][1]};
            adj_gather<<<nblk(b.n), 256, 0, s>>>(ga);
        }
        if ((rc = check_hip(hipGetLastError(), "adjoint kernels"))) return rc;
        double *t = lam; lam = other; other = t;
    }
    return LSM_OK;
}

}  // namespace lsm
