// Pseudo-time reinitialisation of the level-set function.
//
// Sussman/ENO2 path (reinitialisation_test2 + calc_NormGrad_d): each RK stage
// is one kernel that evaluates |grad d| at a node and applies the update
// there (s(phi0), delta(d) and the Sussman-Fatemi lambda are pointwise).
// Interior nodes run in 64 x 4 column tiles marching up k with the stencil of
// d in LDS and registers, so a stage reads d, phi0 and the nine metric terms
// once; boundary nodes (one-sided, edge and vertex formulas) run in extra
// workgroups of the same launch.  The stage-1 |grad d| is the convergence
// mask of the iteration (d^(1) = d^n), kept as one byte per node for stages 2
// and 3.  The exit test of the next
// iteration (max/min of |grad d| over the block) is reduced inside stage 3
// with wavefront shuffles and 64-bit atomics; a one-thread kernel turns it
// into a device flag, so all RNitermax iterations are queued without a host
// round trip and iterations after convergence return at once.
//
// Godunov path (reinitialization_benchmark + calc_sign_ini / calc_RH_RN /
// calc_lambda_RN): one kernel per RK stage; the lambda correction is fused
// into the last stage.
#include <stdlib.h>

#include "lsm_internal.hpp"
#include "lsm_reinit_common.hpp"
#include "lsm_multi.hpp"

namespace lsm {

using rn::Geo;
using rn::make_geo;
using rn::node_of;
using rn::normgrad_node;

namespace {

struct Ctl {
    unsigned long long maxbits, minbits;
    int active, iters;
};

constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;

__global__ __launch_bounds__(256) void normgrad_kernel(Geo g, const double *__restrict__ d,
                                                        const double *__restrict__ phi0, int order_bc,
                                                        int eno_bc, double *__restrict__ ng)
{
    const long n = (long)g.ni * g.nj * g.nk;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    const double p0s = phi0[(long)(g.ni - 2) + g.sj * (g.nj - 2) + g.sk * (g.nk - 2)];
    ng[l] = normgrad_node(g, d, phi0, p0s, order_bc, eno_bc != 0, i, j, k, l);
}

struct CtlArgs {
    Ctl *ctl;
    double eps;
    int32_t *iters_out;
};

struct KInitCtl {
    using Args = CtlArgs;
    static constexpr int kThreads = 1;
    __device__ __forceinline__ static void run(const Args &a, int, int)
    {
        a.ctl->maxbits = 0ull;
        a.ctl->minbits = kInfBits;
        a.ctl->active = 1;
        a.ctl->iters = 0;
        if (a.iters_out) *a.iters_out = 0;
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ void reinit_init_ctl(Ctl *ctl, int32_t *iters_out) { KInitCtl::run(CtlArgs{ctl, 0.0, iters_out}, 0, 1); }

void KInitCtl::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { reinit_init_ctl<<<g, b, 0, s>>>(a.ctl, a.iters_out); }

// exit test at the top of a pseudo-time iteration (reinitialisation_test2.F90:100-107)
struct KCheck {
    using Args = CtlArgs;
    static constexpr int kThreads = 1;
    __device__ __forceinline__ static void run(const Args &a, int, int)
    {
        Ctl *ctl = a.ctl;
        if (ctl->active) {
            const double mx = __longlong_as_double((long long)ctl->maxbits);
            const double mn = __longlong_as_double((long long)ctl->minbits);
            if (fabs(mx - 1.0) < a.eps && fabs(mn - 1.0) < a.eps) ctl->active = 0;
            else ctl->iters += 1;
        }
        ctl->maxbits = 0ull;
        ctl->minbits = kInfBits;
        if (a.iters_out) *a.iters_out = ctl->iters;
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ void reinit_check(Ctl *ctl, double eps, int32_t *iters_out) { KCheck::run(CtlArgs{ctl, eps, iters_out}, 0, 1); }

void KCheck::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { reinit_check<<<g, b, 0, s>>>(a.ctl, a.eps, a.iters_out); }

// the metric tensor of every node (rn::metric_tensor), component c at gt + c * n
struct MetricArgs {
    const double *m[9];
    long n;
    double *gt;
};

struct KMetric {
    using Args = MetricArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long l = (long)bid * 256 + threadIdx.x;
        if (l >= a.n) return;
        double m[9], G[6];
#pragma unroll
        for (int c = 0; c < 9; ++c) m[c] = __builtin_nontemporal_load(a.m[c] + l);
        rn::metric_tensor(m, G);
#pragma unroll
        for (int c = 0; c < 6; ++c) a.gt[c * a.n + l] = G[c];
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void metric_tensor_kernel(MetricArgs a) { KMetric::run(a, (int)blockIdx.x, (int)gridDim.x); }

void KMetric::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { metric_tensor_kernel<<<g, b, 0, s>>>(a); }

struct StageP {
    double c1, c2, c3, dt, epslsm, eps_r;
    int order_bc, eno_bc;
};

// ---------------------------------------------------------------------------
// One RK stage of reinitialisation_test2.F90:137-181.  Interior nodes (no index on a face) run in
// 64 (i) x 4 (j) column tiles marching up a chunk of k planes: d of the tile's
// current plane with its two-node i/j halo sits in LDS, d(k-2..k+2) of the
// thread's own column in registers, so d is read from HBM once per stage.
// The boundary shell (every node with an index on a face, whose one-sided
// and edge/vertex formulas live in normgrad_node_g) is handled by extra
// workgroups of the same launch, one node per thread, except the i-face
// nodes off the edges, which the tiles take; the two parts write disjoint
// nodes and read only din, so they need no ordering.  |grad d| comes from the
// metric tensor everywhere.
// ---------------------------------------------------------------------------
constexpr int RTX = 64, RTY = 4, RH = 2;

struct RTArgs {
    Geo g;
    StageP p;
    const double *din, *dn;
    double *dout;
    const double *phi0;
    const double *gt;   // metric tensor of the interior nodes: component c at gt + c * gs
    long gs;
    uint8_t *conv;
    Ctl *ctl;
    // tape of the stage for the adjoint (TAPE kernels): per node the
    // linearisation of its update (-c3 dtau s(phi0) x d|grad d|/d(D_a d) x
    // spacing per direction, branch codes, -c3 dtau (|grad d| - 1) in the band)
    double *tq[3];
    uint16_t *tcode;
    double *tT;
    int tiles_i, tiles_j, kc, ntiles, grid_tiles;   // interior tiles (grid_tiles: padded to 8)
    long nI, nJ, nK;                                 // shell nodes on i, j and k faces
    int iface;   // the i-face nodes off the edges run in the tiles (the shell keeps the i-face edges)
};

// the stage update at one node, given |grad d| (reinitialisation_test2.F90:137-181);
// s = sign_phi0(phi0) and del = deltaf(d) are returned for the tape
template <int STAGE>
__device__ __forceinline__ double reinit_update(const StageP &p, double ngv, double dv, double dnv,
                                                double p0, bool upd, double &s, double &del)
{
    s = 0.0;
    del = 0.0;
    if (!upd) return dv;
    s = sign_phi0(p0, p.epslsm);
    del = deltaf(dv, p.epslsm);
    return sussman_update(p.c1, p.c2, p.c3 * p.dt, ngv, dv, dnv, s, del);
}

// block-wide max/min (NW wavefronts, linear id tid) into ctl
template <int NW>
__device__ __forceinline__ void reduce_maxmin_t(double vmax, double vmin, int tid, Ctl *ctl)
{
    for (int off = 32; off > 0; off >>= 1) {
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
        vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
    }
    __shared__ double smax[NW], smin[NW];
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) { smax[wv] = vmax; smin[wv] = vmin; }
    __syncthreads();
    if (tid == 0) {
        double a = smax[0], b = smin[0];
        for (int q = 1; q < NW; ++q) { a = fmax(a, smax[q]); b = fmin(b, smin[q]); }
        atomicMax(&ctl->maxbits, (unsigned long long)__double_as_longlong(a));
        atomicMin(&ctl->minbits, (unsigned long long)__double_as_longlong(b));
    }
}

// the tape entry of one node: w are the direction weights already divided by
// |grad d| and scaled by the spacing, br the branch codes.  The weights are
// written at every node (zero where the node's update does not depend on its
// stencil), so every wavefront stores whole 128-byte lines.
__device__ __forceinline__ void tape_write(const RTArgs &a, long l, bool upd, double ngv, double p0, double s,
                                           double del, const double w[3], int br)
{
    uint16_t code = 0;
    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
    if (upd) {
        if (fabs((del * del) * ngv) < kEpsSims) {   // lambda = 0: rhs = s (|grad d| - 1)
            const double c3dt = a.p.c3 * a.p.dt;
            if (fabs(p0) <= a.p.epslsm) { a.tT[l] = -c3dt * (ngv - 1.0); code |= rn::kHasT; }
            const double f = -c3dt * s;
            if (f != 0.0 && ngv > 0.0) {
                q0 = f * w[0]; q1 = f * w[1]; q2 = f * w[2];
                code |= (uint16_t)(rn::kLive | br);
            }
        }
    }
    // streamed: the tape is read back only by the reverse sweep
    __builtin_nontemporal_store(q0, a.tq[0] + l);
    __builtin_nontemporal_store(q1, a.tq[1] + l);
    __builtin_nontemporal_store(q2, a.tq[2] + l);
    __builtin_nontemporal_store(code, a.tcode + l);
}

// STAGE 0: only the max/min of |grad din| (reinitialisation_test2.F90:96-101).
// TY: tile rows (one wavefront per row)
template <int STAGE, bool TAPE = false, int TY = RTY>
struct KReinit {
    using Args = RTArgs;
    static constexpr int kThreads = RTX * TY;
    __device__ __forceinline__ static void run(const Args &a, int bid, int nb);
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

// the plain first stage is held to 80 VGPRs (6 waves per SIMD)
template <int STAGE, bool TAPE = false, int TY = RTY>
__global__ __launch_bounds__(RTX * TY) __attribute__((amdgpu_waves_per_eu(STAGE == 1 && !TAPE ? 6 : 1)))
void reinit_stage_tiled(RTArgs a)
{
    KReinit<STAGE, TAPE, TY>::run(a, (int)blockIdx.x, (int)gridDim.x);
}

template <int STAGE, bool TAPE, int TY>
void KReinit<STAGE, TAPE, TY>::one(unsigned g, dim3 b, hipStream_t s, const Args &a)
{
    reinit_stage_tiled<STAGE, TAPE, TY><<<g, b, 0, s>>>(a);
}

template <int STAGE, bool TAPE, int TY>
__device__ __forceinline__ void KReinit<STAGE, TAPE, TY>::run(const RTArgs &a, int bid, int)
{
    if (!a.ctl->active) return;
    const Geo &g = a.g;
    const int tid = threadIdx.y * RTX + threadIdx.x;
    const double inf = __longlong_as_double((long long)kInfBits);
    double vmax = 0.0, vmin = inf;

    if (bid >= a.grid_tiles) {
        // ---- boundary shell: one node per thread, generic formulas ----
        const long t = (long)(bid - a.grid_tiles) * (RTX * TY) + tid;
        int i = 0, j = 0, k = 0;
        bool ok = true;
        if (t < a.nI && a.iface) {
            // the edges of the two i faces: j = 1 or nj (every k), then k = 1 or nk
            const long per = 2L * g.nk + 2L * (g.nj - 2), u = t % per;
            i = (t < per) ? 0 : g.ni - 1;
            if (u < 2L * g.nk) { j = (u < g.nk) ? 0 : g.nj - 1; k = (int)(u % g.nk); }
            else { const long v = u - 2L * g.nk; j = 1 + (int)(v % (g.nj - 2)); k = (v < g.nj - 2) ? 0 : g.nk - 1; }
        } else if (t < a.nI) {
            const long per = (long)g.nj * g.nk;
            j = (int)((t % per) % g.nj); k = (int)((t % per) / g.nj);
            i = (t < per) ? 0 : g.ni - 1;
        } else if (t < a.nI + a.nJ) {
            const long u = t - a.nI, per = (long)(g.ni - 2) * g.nk;
            i = 1 + (int)((u % per) % (g.ni - 2)); k = (int)((u % per) / (g.ni - 2));
            j = (u < per) ? 0 : g.nj - 1;
        } else if (t < a.nI + a.nJ + a.nK) {
            const long u = t - a.nI - a.nJ, per = (long)(g.ni - 2) * (g.nj - 2);
            i = 1 + (int)((u % per) % (g.ni - 2)); j = 1 + (int)((u % per) / (g.ni - 2));
            k = (u < per) ? 0 : g.nk - 1;
        } else {
            ok = false;
        }
        if (ok) {
            const long l = (long)i + g.sj * j + g.sk * k;
            const double p0s = a.phi0[(long)(g.ni - 2) + g.sj * (g.nj - 2) + g.sk * (g.nk - 2)];
            double tw[3];
            int tbr = 0;
            const double ngv = TAPE ? rn::normgrad_branch(g, a.din, a.phi0, a.gt, a.gs, p0s, a.p.order_bc,
                                                          a.p.eno_bc != 0, i, j, k, l, tw, tbr)
                                    : rn::normgrad_node_g(g, a.din, a.phi0, a.gt, a.gs, p0s, a.p.order_bc,
                                                          a.p.eno_bc != 0, i, j, k, l);
            if (STAGE > 0) {
                bool upd;
                if (STAGE == 1) { upd = ngv - 1.0 > a.p.eps_r; a.conv[l] = upd ? 1 : 0; }
                else upd = a.conv[l] != 0;
                const double dv = a.din[l];
                const double dnv = (STAGE == 1) ? dv : a.dn[l];
                double sv, del;
                a.dout[l] = reinit_update<STAGE>(a.p, ngv, dv, dnv, a.phi0[l], upd, sv, del);
                if constexpr (TAPE) tape_write(a, l, upd, ngv, a.phi0[l], sv, del, tw, tbr);
            }
            vmax = ngv; vmin = ngv;
        }
        if (STAGE == 3 || STAGE == 0) reduce_maxmin_t<TY>(vmax, vmin, tid, a.ctl);
        return;
    }

    // ---- interior tiles ----
    int tile = bid;
    if ((a.grid_tiles & 7) == 0) tile = (bid & 7) * (a.grid_tiles >> 3) + (bid >> 3);
    const bool live = tile < a.ntiles;
    const int tile_c = live ? tile : 0;
    const int ti = tile_c % a.tiles_i;
    const int t2 = tile_c / a.tiles_i;
    const int tj = t2 % a.tiles_j;
    const int tk = t2 / a.tiles_j;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = ti * RTX + tx, j = tj * TY + ty;
    const int kb = 1 + tk * a.kc;
    const int ke = live ? min(kb + a.kc - 1, g.nk - 2) : kb - 1;

    __shared__ double sm[2][TY + 2 * RH][RTX + 2 * RH];

    const int ic = min(i, g.ni - 1), jc = min(j, g.nj - 1);
    const long col = (long)ic + g.sj * jc;
    const bool has_jh = ty < 4;
    const int rr = (ty < 2) ? ty : ty - 2 + TY + RH;
    const int jh = min(max(tj * TY - RH + rr, 0), g.nj - 1);
    const long col_jh = (long)ic + g.sj * jh;
    const int lin = tid;
    const bool has_ih = lin < 4 * TY;
    const int hr = (lin >> 2) % TY, hc4 = lin & 3;
    const int hcol = (hc4 < 2) ? hc4 : hc4 + RTX;
    const int hi_i = min(max(ti * RTX - RH + hcol, 0), g.ni - 1);
    const int hi_j = min(tj * TY + hr, g.nj - 1);
    const long col_ih = (long)hi_i + g.sj * hi_j;
    auto kcl = [&](int kk) -> long { return g.sk * (long)min(max(kk, 0), g.nk - 1); };
    // Nodes on an i face off its edges run in the tiles as well, with the
    // one-sided i formula of the shell (calc_NormGrad_d.F90:187-367) from the
    // tile's row in LDS and the ENO2 upwind sign of phizero at the last
    // interior node: their cache lines are the ones the tile loads anyway,
    // where one node per thread in the shell took a line per node and field.
    const bool fi = a.iface && live && (i == 0 || i == g.ni - 1) && j >= 1 && j <= g.nj - 2;
    const bool act = (live && i >= 1 && i <= g.ni - 2 && j >= 1 && j <= g.nj - 2) || fi;
    const bool ilo = i == 1, ihi = i == g.ni - 2, jlo = j == 1, jhi = j == g.nj - 2;
    const int fdir = (i == 0) ? 1 : -1;
    const double p0s = a.iface ? a.phi0[(long)(g.ni - 2) + g.sj * (g.nj - 2) + g.sk * (g.nk - 2)] : 0.0;
    const double *__restrict__ d = a.din;

    double q0 = d[col + kcl(kb - 2)], q1 = d[col + kcl(kb - 1)];
    double q2 = d[col + kcl(kb)], q3 = d[col + kcl(kb + 1)];
    double nq4 = d[col + kcl(kb + 2)];
    double njh = has_jh ? d[col_jh + kcl(kb)] : 0.0;
    double nih = has_ih ? d[col_ih + kcl(kb)] : 0.0;
    int buf = 0;
    for (int k = kb; k <= ke; ++k) {
        const double q4 = nq4;
        sm[buf][ty + RH][tx + RH] = q2;
        if (has_jh) sm[buf][rr][tx + RH] = njh;
        if (has_ih) sm[buf][hr + RH][hcol] = nih;
        __syncthreads();
        if (k < ke) {
            const long kn = kcl(k + 1);
            nq4 = d[col + kcl(k + 3)];
            if (has_jh) njh = d[col_jh + kn];
            if (has_ih) nih = d[col_ih + kn];
        }
        if (act) {
            const long l = col + g.sk * (long)k;
            const double p0 = a.phi0[l];
            const double pe = fi ? p0s : p0;   // ENO2 upwind sign
            double G[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) G[c] = __builtin_nontemporal_load(a.gt + c * a.gs + l);
            double dnv = q2;
            uint8_t cv = 0;
            if (STAGE > 1) { dnv = __builtin_nontemporal_load(a.dn + l); cv = a.conv[l]; }
            const double (*pl)[RTX + 2 * RH] = sm[buf];
            const int cx = tx + RH, cy = ty + RH;
            // calc_NormGrad_d.F90:70-163 (ENO2 in each direction, upwind side by phi0)
            double g0;
            if (fi) g0 = g.dd[0] * biased(q2, pl[cy][cx + fdir], pl[cy][cx + 2 * fdir], pl[cy][cx + 3 * fdir],
                                          a.p.order_bc, fdir, a.p.eno_bc != 0);
            else g0 = g.dd[0] * eno2(pe, ilo ? 0.0 : pl[cy][cx - 2], pl[cy][cx - 1], q2, pl[cy][cx + 1],
                                     ihi ? 0.0 : pl[cy][cx + 2], ilo, ihi);
            const double g1 = g.dd[1] * eno2(pe, jlo ? 0.0 : pl[cy - 2][cx], pl[cy - 1][cx], q2,
                                             pl[cy + 1][cx], jhi ? 0.0 : pl[cy + 2][cx], jlo, jhi);
            const bool klo = k == 1, khi = k == g.nk - 2;
            const double g2 = g.dd[2] * eno2(pe, klo ? 0.0 : q0, q1, q2, q3, khi ? 0.0 : q4, klo, khi);
            double t[3];
            const double ngv = rn::tensor_norm(G, g0, g1, g2, t);
            if (STAGE > 0) {
                bool upd;
                if (STAGE == 1) { upd = ngv - 1.0 > a.p.eps_r; a.conv[l] = upd ? 1 : 0; }
                else upd = cv != 0;
                double sv, del;
                __builtin_nontemporal_store(reinit_update<STAGE>(a.p, ngv, q2, dnv, p0, upd, sv, del), a.dout + l);
                if constexpr (TAPE) {
                    int b0 = 0;
                    if (!fi) b0 = eno2_branch(pe, ilo ? 0.0 : pl[cy][cx - 2], pl[cy][cx - 1], q2, pl[cy][cx + 1],
                                              ihi ? 0.0 : pl[cy][cx + 2], ilo, ihi);
                    else if (a.p.order_bc == 2 && a.p.eno_bc)
                        b0 = eno2_boundary_branch(q2, pl[cy][cx + fdir], pl[cy][cx + 2 * fdir], fdir);
                    const int br = b0 |
                                   (eno2_branch(pe, jlo ? 0.0 : pl[cy - 2][cx], pl[cy - 1][cx], q2, pl[cy + 1][cx],
                                                jhi ? 0.0 : pl[cy + 2][cx], jlo, jhi) << 4) |
                                   (eno2_branch(pe, klo ? 0.0 : q0, q1, q2, q3, khi ? 0.0 : q4, klo, khi) << 8);
                    // 1 / |grad d|: hardware reciprocal and two Newton steps
                    double inv = __builtin_amdgcn_rcp(ngv);
                    inv = fma(inv, fma(-ngv, inv, 1.0), inv);
                    inv = fma(inv, fma(-ngv, inv, 1.0), inv);
                    inv = ngv > 0.0 ? inv : 0.0;
                    double tw[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) tw[d] = t[d] * inv * g.dd[d];
                    tape_write(a, l, upd, ngv, p0, sv, del, tw, br);
                }
            }
            if (STAGE == 3 || STAGE == 0) { vmax = fmax(vmax, ngv); vmin = fmin(vmin, ngv); }
        }
        q0 = q1; q1 = q2; q2 = q3; q3 = q4;
        buf ^= 1;
    }
    if (STAGE == 3 || STAGE == 0) reduce_maxmin_t<TY>(vmax, vmin, tid, a.ctl);
}

// ---------------------------------------------------------------------------
// Godunov path
// ---------------------------------------------------------------------------

// one-sided differences with the boundary copies of calc_sign_ini.F90:41-176,
// then the Godunov |grad| pair of calc_sign_ini.F90:204-242
__device__ void godunov_norm(const Geo &g, const double *__restrict__ f, long l, int i, int j, int k,
                             double &ngp, double &ngm)
{
    const int n[3] = {g.ni, g.nj, g.nk};
    const int idx[3] = {i, j, k};
    const long st[3] = {1, g.sj, g.sk};
    double mas[3], men[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long s = st[a];
        if (idx[a] == 0) {
            mas[a] = g.dd[a] * (f[l + s] - f[l]);
            men[a] = mas[a];
        } else if (idx[a] == n[a] - 1) {
            men[a] = g.dd[a] * (f[l] - f[l - s]);
            mas[a] = men[a];
        } else {
            mas[a] = g.dd[a] * (f[l + s] - f[l]);
            men[a] = g.dd[a] * (f[l] - f[l - s]);
        }
    }
    double bm[3], bn[3], fm[3], fn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double db = (g.m[c][l] * men[0] + g.m[3 + c][l] * men[1]) + g.m[6 + c][l] * men[2];
        const double df = (g.m[c][l] * mas[0] + g.m[3 + c][l] * mas[1]) + g.m[6 + c][l] * mas[2];
        double a = db / 2.0;
        bm[c] = a + fabs(a); bn[c] = a - fabs(a);
        a = df / 2.0;
        fm[c] = a + fabs(a); fn[c] = a - fabs(a);
    }
    auto mx = [](double p, double q) { return p > q ? p : q; };
    ngp = sqrt((mx(bm[0] * bm[0], fn[0] * fn[0]) + mx(bm[1] * bm[1], fn[1] * fn[1]))
               + mx(bm[2] * bm[2], fn[2] * fn[2]));
    ngm = sqrt((mx(bn[0] * bn[0], fm[0] * fm[0]) + mx(bn[1] * bn[1], fm[1] * fm[1]))
               + mx(bn[2] * bn[2], fm[2] * fm[2]));
}

// calc_sign_ini.F90:199-270
__global__ __launch_bounds__(256) void sign_ini_kernel(Geo g, const double *__restrict__ phi0, double eps,
                                                        int m_sphi0, double *__restrict__ sgn,
                                                        double *__restrict__ ng)
{
    const long n = (long)g.ni * g.nj * g.nk;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    double gp, gm;
    godunov_norm(g, phi0, l, i, j, k, gp, gm);
    const double p = phi0[l];
    const double gv = (p > 0.0) ? gp : ((p < 0.0) ? gm : 1.0);
    double s;
    if (m_sphi0 == 1) {
        if (p >= eps) s = 1.0;
        else if (p <= (-1.0) * eps) s = -1.0;
        else s = p / eps - (1.0 / kPi) * sin((kPi * p) / eps);
    } else {
        s = p / sqrt(p * p + (gv * eps) * (gv * eps));
    }
    sgn[l] = s;
    ng[l] = gv;
}

// calc_RH_RN.F90:199-240
__device__ __forceinline__ double rh_rn_node(const Geo &g, const double *__restrict__ d, double s,
                                             long l, int i, int j, int k)
{
    double gp, gm;
    godunov_norm(g, d, l, i, j, k, gp, gm);
    const double gv = (s > 0.0) ? gp : ((s < 0.0) ? gm : 1.0);
    return s * (gv - 1.0);
}

__global__ __launch_bounds__(256) void rh_rn_kernel(Geo g, const double *__restrict__ d,
                                                     const double *__restrict__ sgn, double *__restrict__ rh)
{
    const long n = (long)g.ni * g.nj * g.nk;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    rh[l] = rh_rn_node(g, d, sgn[l], l, i, j, k);
}

// calc_lambda_RN.F90:42-63 (method 0)
__device__ __forceinline__ double lambda_node(double p0, double ng0, double d, double dt, double eps,
                                              double dc, double de, double dz)
{
    const double vol = ((1.0 / dc) * (1.0 / de)) * (1.0 / dz);
    const double Lop = deltaf(p0, eps) * ((1.0 / dt) * (d - p0));
    const double vn = vol * Lop;
    const double del = deltaf(p0, eps);
    const double vd = (vol * (del * del)) * ng0;
    return (vd == 0.0) ? 0.0 : -(vn / vd);
}

__global__ __launch_bounds__(256) void lambda_kernel(long n, const double *__restrict__ phi0,
                                                      const double *__restrict__ ng0,
                                                      const double *__restrict__ d, double dt, double eps,
                                                      double dc, double de, double dz,
                                                      double *__restrict__ lam)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    lam[l] = lambda_node(phi0[l], ng0[l], d[l], dt, eps, dc, de, dz);
}

// calc_lambda_RN.F90:65-120 (sussman_correction_method = 1): the pointwise
// numerator and denominator of every node, filtered by testfilter below
__global__ __launch_bounds__(256) void lambda_parts_kernel(long n, const double *__restrict__ phi0,
                                                           const double *__restrict__ ng0,
                                                           const double *__restrict__ d, double dt, double eps,
                                                           double *__restrict__ num, double *__restrict__ den)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const double del = deltaf(phi0[l], eps);
    num[l] = del * ((1.0 / dt) * (d[l] - phi0[l]));
    den[l] = (del * del) * ng0[l];
}

// testfilter.F90:46-121 on the numerator and the denominator, then the
// quotient of calc_lambda_RN.F90:94-118.  At the nodes off the block's faces
// the 27-node average with weights 64 / 16 / 4 / 1 (the node, face, edge and
// corner neighbours) divided by aj there, in the reference's order (both
// filters share the weights and their sum); the face nodes keep the
// pointwise values.  lambda = 0 where the filtered denominator is within 1E-10
// (a default-real literal in the reference) of 0.
__global__ __launch_bounds__(256) void lambda_filter_kernel(Geo g, const double *__restrict__ aj,
                                                            const double *__restrict__ num,
                                                            const double *__restrict__ den,
                                                            double *__restrict__ lam)
{
    const long n = (long)g.ni * g.nj * g.nk;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    double fn = num[l], fd = den[l];
    if (i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1 && k > 0 && k < g.nk - 1) {
        // testfilter.F90:58-84 order: node, +-i, +-j, +-k, the i-j, i-k and
        // j-k edges, then the corners
        constexpr signed char off[27][3] = {
            {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
            {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
            {-1, 0, -1}, {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},   {1, 1, -1},
            {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
        double sn = 0.0, sd = 0.0, sv = 0.0;
#pragma unroll
        for (int m = 0; m < 27; ++m) {
            const int nz = (off[m][0] != 0) + (off[m][1] != 0) + (off[m][2] != 0);
            const double c = nz == 0 ? 64.0 : nz == 1 ? 16.0 : nz == 2 ? 4.0 : 1.0;
            const long q = l + off[m][0] + g.sj * off[m][1] + g.sk * off[m][2];
            const double w = c / aj[q];
            sn = sn + w * num[q];
            sd = sd + w * den[q];
            sv = sv + w;
        }
        fn = sn / sv;
        fd = sd / sv;
    }
    lam[l] = (fabs(fd) <= (double)1e-10f) ? 0.0 : -(fn / fd);
}

// the correction of reinitialization_benchmark.F90:166-181 with a lambda field
__global__ __launch_bounds__(256) void lambda_apply_kernel(long n, const double *__restrict__ v,
                                                           const double *__restrict__ lam,
                                                           const double *__restrict__ phi0,
                                                           const double *__restrict__ ng0, double dt, double eps,
                                                           double *__restrict__ out)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    out[l] = v[l] + ((dt * lam[l]) * deltaf(phi0[l], eps)) * ng0[l];
}

// one RK stage (or the Euler step) of reinitialization_benchmark.F90:96-158;
// LAMBDA fuses the correction of reinitialization_benchmark.F90:166-181
template <bool LAMBDA>
__global__ __launch_bounds__(256) void godunov_stage(Geo g, const double *__restrict__ din,
                                                      const double *sn, double *dout,
                                                      const double *__restrict__ sgn,
                                                      const double *__restrict__ phi0,
                                                      const double *__restrict__ ng0, double c1,
                                                      double c2, double c3, double dt, double eps)
{
    const long n = (long)g.ni * g.nj * g.nk;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    const double rh = rh_rn_node(g, din, sgn[l], l, i, j, k);
    double v;
    if (c3 < 0.0) v = sn[l] - dt * rh;                         // forward Euler
    else v = (c1 * sn[l] + c2 * din[l]) - (c3 * dt) * rh;
    if (LAMBDA) {
        const double lam = lambda_node(phi0[l], ng0[l], v, dt, eps, g.dd[0], g.dd[1], g.dd[2]);
        v = v + ((dt * lam) * deltaf(phi0[l], eps)) * ng0[l];
    }
    dout[l] = v;
}

inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

size_t fb(const Blk &b) { return ((size_t)b.n * sizeof(double) + 255) & ~(size_t)255; }

}  // namespace

size_t reinit_godunov_work(const Blk &b) { return 7 * fb(b) + 1024; }

size_t lambda_simpson_work(const Blk &b) { return 2 * fb(b) + 512; }

int launch_metric_tensor(const double *const m[9], long n, double *gt, hipStream_t s)
{
    MetricArgs a{};
    for (int c = 0; c < 9; ++c) a.m[c] = m[c];
    a.n = n;
    a.gt = gt;
    return klaunch<KMetric>(nb(n), dim3(256), s, a, "metric_tensor_kernel");
}

int launch_metric_tensor(const Blk &b, double *gt, hipStream_t s)
{
    const double *m[9] = {b.csi0, b.csi1, b.csi2, b.eta0, b.eta1, b.eta2, b.zet0, b.zet1, b.zet2};
    return launch_metric_tensor(m, b.n, gt, s);
}

// the block's metric tensor: the caller's, or computed into the workspace
static int block_tensor(const Blk &b, Carver &w, hipStream_t s, const double **gt)
{
    if (b.gten) { *gt = b.gten; return LSM_OK; }
    double *t = (double *)w.take(6 * (size_t)b.n * sizeof(double));
    if (!t) return fail(LSM_ERR_WORKSPACE, "reinit workspace too small (metric tensor)");
    *gt = t;
    return launch_metric_tensor(b, t, s);
}

// workspace of block_tensor
static size_t tensor_work(const Blk &b) { return b.gten ? 0 : ((6 * (size_t)b.n * sizeof(double) + 255) & ~(size_t)255); }

size_t reinit_eno2_work(const Blk &b) { return 2 * fb(b) + (((size_t)b.n + 255) & ~(size_t)255) + 1024 + tensor_work(b); }

int launch_normgrad(const Blk &b, const double *d, const double *phi0, int order_bc, int eno_bc,
                    double *ng, hipStream_t s)
{
    normgrad_kernel<<<nb(b.n), 256, 0, s>>>(make_geo(b), d, phi0, order_bc, eno_bc, ng);
    return check_hip(hipGetLastError(), "normgrad_kernel");
}

namespace {

const double kRkC[3][3] = {{1.0, 0.0, 1.0},
                           {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                           {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

// launch shape of the tiled stage kernel: interior tiles of 64 x 4 nodes x kc
// planes (padded to a multiple of 8 workgroups) plus the boundary-shell blocks
struct RTShape {
    RTArgs a;
    unsigned grid;
    int ty;   // tile rows
};

// ty: tile rows, 4 for the plain stages, 8 for the taped ones (whose wider
// writes make the taller tile's smaller j halo pay: 1.6 % on the step adjoint
// at 256^3, 0.5 % slower for the plain stages)
RTShape rt_shape(const Blk &b, const LsmReinitParams &p, const double *phi0, uint8_t *conv, Ctl *ctl,
                 const double *gt, int ty = RTY)
{
    RTShape r{};
    RTArgs &ra = r.a;
    ra.g = make_geo(b); ra.phi0 = phi0; ra.conv = conv; ra.ctl = ctl;
    ra.gt = gt; ra.gs = b.n;
    r.ty = ty == 8 ? 8 : RTY;
    ra.tiles_i = (b.e.ni + RTX - 1) / RTX;
    ra.tiles_j = (b.e.nj + r.ty - 1) / r.ty;
    const int nk_int = b.e.nk - 2;
    const int tij = ra.tiles_i * ra.tiles_j;
    // about 4096 tiles (16 planes per tile at 256^3: the k halo of a tile is
    // 4 planes of d per chunk)
    int kc = (nk_int + (4096 + tij - 1) / tij - 1) / ((4096 + tij - 1) / tij);
    if (kc < 8) kc = 8;
    if (const char *e = getenv("LSM_REINIT_KCHUNK")) { int v = atoi(e); if (v > 0) kc = v; }
    ra.kc = kc;
    ra.ntiles = tij * ((nk_int + kc - 1) / kc);
    ra.grid_tiles = (ra.ntiles + 7) / 8 * 8;
    // i-face nodes off the edges in the tiles where the one-sided i stencil of
    // i = ni stays inside its tile's row and halo (not in lane 0 of a tile)
    ra.iface = b.e.ni >= 5 && b.e.nj >= 3 && b.e.nk >= 3 && (b.e.ni - 1) % RTX != 0;
    ra.nI = ra.iface ? 2L * (2L * b.e.nk + 2L * (b.e.nj - 2)) : 2L * b.e.nj * b.e.nk;
    ra.nJ = 2L * (b.e.ni - 2) * b.e.nk;
    ra.nK = 2L * (b.e.ni - 2) * (b.e.nj - 2);
    const long wg = (long)RTX * r.ty;
    const unsigned shell_blocks = (unsigned)((ra.nI + ra.nJ + ra.nK + wg - 1) / wg);
    r.grid = (unsigned)ra.grid_tiles + shell_blocks;
    ra.p = StageP{0, 0, 0, p.rn_timestep, p.epslsm, p.eps_reinit, p.order_bc, p.eno_bc};
    return r;
}

template <int TY>
int rt_launch_ty(RTShape &r, int stage, hipStream_t s, bool tape)
{
    const dim3 rblk(RTX, TY, 1);
    const char *w = "reinit_stage_tiled";
    switch (stage) {
    case 0: return klaunch<KReinit<0, false, TY>>(r.grid, rblk, s, r.a, w);
    case 1: return tape ? klaunch<KReinit<1, true, TY>>(r.grid, rblk, s, r.a, w)
                        : klaunch<KReinit<1, false, TY>>(r.grid, rblk, s, r.a, w);
    case 2: return tape ? klaunch<KReinit<2, true, TY>>(r.grid, rblk, s, r.a, w)
                        : klaunch<KReinit<2, false, TY>>(r.grid, rblk, s, r.a, w);
    default: return tape ? klaunch<KReinit<3, true, TY>>(r.grid, rblk, s, r.a, w)
                         : klaunch<KReinit<3, false, TY>>(r.grid, rblk, s, r.a, w);
    }
}

int rt_launch(RTShape &r, int stage, hipStream_t s, bool tape = false)
{
    return r.ty == 8 ? rt_launch_ty<8>(r, stage, s, tape) : rt_launch_ty<RTY>(r, stage, s, tape);
}

int init_ctl(Ctl *ctl, int32_t *iters, hipStream_t s)
{
    return klaunch<KInitCtl>(1, dim3(1), s, CtlArgs{ctl, 0.0, iters}, "reinit_init_ctl");
}

int check_ctl(Ctl *ctl, double eps, int32_t *iters, hipStream_t s)
{
    return klaunch<KCheck>(1, dim3(1), s, CtlArgs{ctl, eps, iters}, "reinit_check");
}

void rt_stage_params(RTShape &r, const LsmReinitParams &p, int st)
{
    r.a.p = StageP{kRkC[st][0], kRkC[st][1], kRkC[st][2], p.rn_timestep, p.epslsm, p.eps_reinit,
                   p.order_bc, p.eno_bc};
}

}  // namespace

int launch_reinit_eno2(const Blk &b, const LsmReinitParams &p, const double *phi0, double *out,
                       int32_t *iters, Carver &w, hipStream_t s)
{
    double *A = (double *)w.take(b.n * sizeof(double));
    double *B = (double *)w.take(b.n * sizeof(double));
    uint8_t *conv = (uint8_t *)w.take((size_t)b.n);
    Ctl *ctl = (Ctl *)w.take(sizeof(Ctl));
    if (!A || !B || !conv || !ctl) return fail(LSM_ERR_WORKSPACE, "reinit workspace too small");
    int rc;
    if ((rc = launch_copy(phi0, out, b.n, s))) return rc;
    if ((rc = init_ctl(ctl, iters, s))) return rc;
    if (p.rn_itermax <= 0) return LSM_OK;
    const double *gt = nullptr;
    if ((rc = block_tensor(b, w, s, &gt))) return rc;
    RTShape r = rt_shape(b, p, phi0, conv, ctl, gt);
    // max/min of |grad phi0| before the first iteration (reinitialisation_test2.F90:96-101)
    r.a.din = phi0; r.a.dn = phi0; r.a.dout = nullptr;
    if ((rc = rt_launch(r, 0, s))) return rc;
    const double *sin_[3] = {out, A, B};
    const double *sdn[3] = {out, out, out};
    double *sout[3] = {A, B, out};
    for (int it = 0; it < p.rn_itermax; ++it) {
        if ((rc = check_ctl(ctl, p.eps_reinit, iters, s))) return rc;
        for (int st = 0; st < 3; ++st) {
            KTimer kt(LSM_KT_REINIT_STAGE, s);
            rt_stage_params(r, p, st);
            r.a.din = sin_[st]; r.a.dn = sdn[st]; r.a.dout = sout[st];
            if ((rc = rt_launch(r, st + 1, s))) return rc;
        }
    }
    return LSM_OK;
}

size_t reinit_trace_work(const Blk &b, int itermax)
{
    const size_t u8 = ((size_t)b.n + 255) & ~(size_t)255;
    const size_t u16 = ((size_t)b.n * sizeof(uint16_t) + 255) & ~(size_t)255;
    return (size_t)itermax * (u8 + 3 * (4 * fb(b) + u16)) + 2 * fb(b) + 2048 + tensor_work(b);
}

int carve_reinit_trace(const Blk &b, const LsmReinitParams &p, ReinitTape &t, Carver &w, void **ctl_out,
                       double **A, double **B)
{
    const int nit = p.rn_itermax;
    if (nit < 0 || nit > kMaxTraceIters)
        return fail(LSM_ERR_UNSUPPORTED, "rn_itermax = %d outside 0..%d for the adjoint", nit, kMaxTraceIters);
    const size_t f = (size_t)b.n * sizeof(double);
    t.nit = nit;
    for (int it = 0; it < nit; ++it) {
        t.conv[it] = (uint8_t *)w.take((size_t)b.n);
        if (!t.conv[it]) return fail(LSM_ERR_WORKSPACE, "reinit trace workspace too small");
        for (int st = 0; st < 3; ++st) {
            for (int d = 0; d < 3; ++d) t.q[it][st][d] = (double *)w.take(f);
            t.T[it][st] = (double *)w.take(f);
            t.code[it][st] = (uint16_t *)w.take((size_t)b.n * sizeof(uint16_t));
            if (!t.q[it][st][0] || !t.q[it][st][1] || !t.q[it][st][2] || !t.T[it][st] || !t.code[it][st])
                return fail(LSM_ERR_WORKSPACE, "reinit trace workspace too small");
        }
    }
    *A = (double *)w.take(f);
    *B = (double *)w.take(f);
    void *ctl = w.take(sizeof(Ctl));
    t.iters = (int32_t *)w.take(sizeof(int32_t));
    if (!*A || !*B || !ctl || !t.iters) return fail(LSM_ERR_WORKSPACE, "reinit trace workspace too small");
    if (ctl_out) *ctl_out = ctl;
    return LSM_OK;
}

int launch_reinit_eno2_trace(const Blk &b, const LsmReinitParams &p, const double *phi0, ReinitTape &t,
                             double *out, Carver &w, hipStream_t s)
{
    void *cv = nullptr;
    double *A = nullptr, *B = nullptr;
    int rc = carve_reinit_trace(b, p, t, w, &cv, &A, &B);
    if (rc) return rc;
    if (!out || out == phi0) return fail(LSM_ERR_ARG, "reinit trace output missing or aliasing phi0");
    Ctl *ctl = (Ctl *)cv;
    if ((rc = launch_copy(phi0, out, b.n, s))) return rc;
    if ((rc = init_ctl(ctl, t.iters, s))) return rc;
    if (t.nit == 0) return LSM_OK;
    const double *gt = nullptr;
    if ((rc = block_tensor(b, w, s, &gt))) return rc;
    RTShape r = rt_shape(b, p, phi0, nullptr, ctl, gt, 8);
    r.a.din = phi0; r.a.dn = phi0; r.a.dout = nullptr;
    if ((rc = rt_launch(r, 0, s))) return rc;
    const double *sin_[3] = {out, A, B};
    double *sout[3] = {A, B, out};
    for (int it = 0; it < t.nit; ++it) {
        if ((rc = check_ctl(ctl, p.eps_reinit, t.iters, s))) return rc;
        r.a.conv = t.conv[it];
        for (int st = 0; st < 3; ++st) {
            rt_stage_params(r, p, st);
            r.a.din = sin_[st]; r.a.dn = out; r.a.dout = sout[st];
            for (int d = 0; d < 3; ++d) r.a.tq[d] = t.q[it][st][d];
            r.a.tcode = t.code[it][st];
            r.a.tT = t.T[it][st];
            if ((rc = rt_launch(r, st + 1, s, true))) return rc;
        }
    }
    return LSM_OK;
}

int launch_sign_ini(const Blk &b, const double *phi0, double eps, int m_sphi0, double *sgn, double *ng,
                    hipStream_t s)
{
    sign_ini_kernel<<<nb(b.n), 256, 0, s>>>(make_geo(b), phi0, eps, m_sphi0, sgn, ng);
    return check_hip(hipGetLastError(), "sign_ini_kernel");
}

int launch_rh_rn(const Blk &b, const double *d, const double *sgn, double *rh, hipStream_t s)
{
    rh_rn_kernel<<<nb(b.n), 256, 0, s>>>(make_geo(b), d, sgn, rh);
    return check_hip(hipGetLastError(), "rh_rn_kernel");
}

int launch_lambda_rn(const Blk &b, const double *phi0, const double *ng0, const double *d, double dt,
                     double eps, double *lam, hipStream_t s)
{
    lambda_kernel<<<nb(b.n), 256, 0, s>>>(b.n, phi0, ng0, d, dt, eps, b.dc, b.de, b.dz, lam);
    return check_hip(hipGetLastError(), "lambda_kernel");
}

int launch_lambda_rn_simpson(const Blk &b, const double *phi0, const double *ng0, const double *d, double dt,
                             double eps, double *lam, Carver &w, hipStream_t s)
{
    const size_t f = b.n * sizeof(double);
    double *num = (double *)w.take(f), *den = (double *)w.take(f);
    if (!num || !den) return fail(LSM_ERR_WORKSPACE, "lambda workspace too small");
    lambda_parts_kernel<<<nb(b.n), 256, 0, s>>>(b.n, phi0, ng0, d, dt, eps, num, den);
    int rc = check_hip(hipGetLastError(), "lambda_parts_kernel");
    if (rc) return rc;
    lambda_filter_kernel<<<nb(b.n), 256, 0, s>>>(make_geo(b), b.aj, num, den, lam);
    return check_hip(hipGetLastError(), "lambda_filter_kernel");
}

int launch_reinit_godunov(const Blk &b, const LsmReinitParams &p, const double *phi0, double *out,
                          Carver &w, hipStream_t s)
{
    const Geo g = make_geo(b);
    const size_t f = b.n * sizeof(double);
    double *sgn = (double *)w.take(f), *ng0 = (double *)w.take(f);
    double *A = (double *)w.take(f), *B = (double *)w.take(f);
    if (!sgn || !ng0 || !A || !B) return fail(LSM_ERR_WORKSPACE, "reinit workspace too small");
    int rc;
    if ((rc = launch_copy(phi0, out, b.n, s))) return rc;
    if ((rc = launch_sign_ini(b, phi0, p.epslsm, p.m_sphi0, sgn, ng0, s))) return rc;
    const double dt = p.rn_timestep, eps = p.epslsm;
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    if (p.lambda_method == 1) {
        // filtered lambda (calc_lambda_RN.F90:65-120): the stages without the
        // fused correction, then lambda over the whole field and the correction
        double *lam = (double *)w.take(f);
        if (!lam) return fail(LSM_ERR_WORKSPACE, "reinit workspace too small");
        for (int it = 0; it < p.rn_itermax; ++it) {
            if (p.m_ti == 1) {
                godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, out, out, A, sgn, phi0, ng0, C[0][0], C[0][1], C[0][2], dt, eps);
                godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, A, out, B, sgn, phi0, ng0, C[1][0], C[1][1], C[1][2], dt, eps);
                godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, B, out, A, sgn, phi0, ng0, C[2][0], C[2][1], C[2][2], dt, eps);
            } else {
                godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, out, out, A, sgn, phi0, ng0, 0.0, 0.0, -1.0, dt, eps);
            }
            if ((rc = check_hip(hipGetLastError(), "godunov_stage"))) return rc;
            Carver lw = w;
            if ((rc = launch_lambda_rn_simpson(b, phi0, ng0, A, dt, eps, lam, lw, s))) return rc;
            lambda_apply_kernel<<<nb(b.n), 256, 0, s>>>(b.n, A, lam, phi0, ng0, dt, eps, out);
            if ((rc = check_hip(hipGetLastError(), "lambda_apply_kernel"))) return rc;
        }
        return LSM_OK;
    }
    for (int it = 0; it < p.rn_itermax; ++it) {
        if (p.m_ti == 1) {
            godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, out, out, A, sgn, phi0, ng0, C[0][0], C[0][1], C[0][2], dt, eps);
            godunov_stage<false><<<nb(b.n), 256, 0, s>>>(g, A, out, B, sgn, phi0, ng0, C[1][0], C[1][1], C[1][2], dt, eps);
            godunov_stage<true><<<nb(b.n), 256, 0, s>>>(g, B, out, out, sgn, phi0, ng0, C[2][0], C[2][1], C[2][2], dt, eps);
        } else {
            godunov_stage<true><<<nb(b.n), 256, 0, s>>>(g, out, out, A, sgn, phi0, ng0, 0.0, 0.0, -1.0, dt, eps);
            if ((rc = launch_copy(A, out, b.n, s))) return rc;
        }
        if ((rc = check_hip(hipGetLastError(), "godunov_stage"))) return rc;
    }
    return LSM_OK;
}

}  // namespace lsm
