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
// The forward iterations are run once more with a tape (lsm_reinit.hip): per
// stage and node the linearisation of the update, q_d = -c3 dtau s(phi0) x
// d|grad d|/d(D_d d) x spacing per direction, a 4-bit code per direction of
// the branch its difference formula took (ENO2 result and second differences,
// or the ENO2Boundary case), and inside the interface band T = -c3 dtau
// (|grad d| - 1).  Per stage (reverse order) one kernel, `radj_tape_gather`,
// forms at each node the sum over the nodes whose stencil holds it (slots:
// offsets -2..2 inside, 0..3 away from a face on a face) of their adjoint x
// q_d x the coefficient rebuilt from the code, plus the linear parts, and
// accumulates adjoint x T for dJ/ds(phi0), turned into dJ/dphi0 at the end.
// Deterministic (no atomics).
#include <stdlib.h>

#include <algorithm>

#include "lsm_internal.hpp"
#include "lsm_reinit_common.hpp"
#include "lsm_multi.hpp"

namespace lsm {

namespace {

using rn::Geo;
using rn::make_geo;
using rn::node_of;

using rn::kLive;
using rn::kHasT;

struct TGatherArgs {
    int ni, nj, nk;
    long sj, sk, n;
    const double *lam;
    const double *q[3];
    const uint16_t *code;
    const double *T;
    const uint8_t *conv;
    const int32_t *iters;
    int it, stage, order_bc, eno_bc;
    double c1, c2;
    double *out;     // adjoint of the stage input
    double *adn;     // adjoint of the iteration input d^n, accumulated over stages 3, 2
    double *adj_s;   // accumulated dJ/ds(phi0)
};

// One reverse stage: at each node the sum over the nodes whose stencil holds
// it of their adjoint x q_d x the coefficient rebuilt from their branch code,
// plus the linear parts.  A workgroup of 64 (i) x GY (j) threads marches up a
// chunk of k planes; per
// plane the products lam q_0 (with a 3-node i halo), lam q_1 (3-node j halo)
// and the branch codes (both halos) sit in LDS, and lam q_2 and the codes of
// planes k-3 .. k+3 of the thread's own column in a register window.
constexpr int GX = 64, GH = 3;

// the coefficient of a face node's one-sided difference (biased_coef) at one
// slot, or of the i = 1, k = 1 edge's fixed second-order stencil; faces only,
// kept out of line
__device__ __noinline__ double face_coef(int br, int order_bc, int dir, bool enobc, int slot, bool edge)
{
    double e[4];
    if (edge) { e[0] = -3.0; e[1] = 4.0; e[2] = -1.0; e[3] = 0.0; }
    else biased_coef(br, order_bc, dir, enobc, e);
    return slot == 0 ? e[0] : slot == 1 ? e[1] : slot == 2 ? e[2] : e[3];
}

struct TGatherLaunch {
    TGatherArgs a;
    int tiles_i, tiles_j, kc;
};

template <int GY>
__device__ __forceinline__ void radj_tape_gather_run(const TGatherArgs &a, int tiles_i, int tiles_j, int kc, int bid,
                                                     int nb);

template <int GY>
struct KTGather {
    using Args = TGatherLaunch;
    static constexpr int kThreads = GX * GY;
    __device__ __forceinline__ static void run(const Args &x, int bid, int nb)
    {
        radj_tape_gather_run<GY>(x.a, x.tiles_i, x.tiles_j, x.kc, bid, nb);
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &x);
};

template <int GY>
__global__ __launch_bounds__(GX * GY) void radj_tape_gather(TGatherLaunch x)
{
    KTGather<GY>::run(x, (int)blockIdx.x, (int)gridDim.x);
}

template <int GY>
void KTGather<GY>::one(unsigned g, dim3 b, hipStream_t s, const Args &x)
{
    radj_tape_gather<GY><<<g, b, 0, s>>>(x);
}

template <int GY>
__device__ __forceinline__ void radj_tape_gather_run(const TGatherArgs &a, int tiles_i, int tiles_j, int kc, int bid,
                                                     int nb)
{
    // j halo cells (2 GH rows x GX columns) and i halo cells (GY rows x 2 GH
    // columns) are loaded by one thread each
    constexpr int NT = GX * GY, NJH = 2 * GH * GX, NIH = 2 * GH * GY;
    static_assert(NJH + NIH <= NT, "halo cells exceed the workgroup");
    __shared__ double sW0[2][GY][GX + 2 * GH];     // lam q_0 of the plane, i halo
    __shared__ double sW1[2][GY + 2 * GH][GX];     // lam q_1 of the plane, j halo
    __shared__ uint16_t sCi[2][GY][GX + 2 * GH];   // branch codes, i halo
    __shared__ uint16_t sCj[2][GY + 2 * GH][GX];   // branch codes, j halo
    __shared__ double s_eno[16][5];                // interior coefficients of every ENO2 branch code
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * GX + tx;
    if (tid < 16) {
        double e[5];
        eno2_coef(tid, e);
        for (int q = 0; q < 5; ++q) s_eno[tid][q] = e[q];
    }
    // consecutive tiles on one XCD (workgroup bid runs on XCD bid % 8), so the
    // halos a tile shares with its neighbours are read from one L2
    const int ntile = tiles_i * tiles_j * ((a.nk + kc - 1) / kc);
    const int b = (nb & 7) == 0 ? (bid & 7) * (nb >> 3) + (bid >> 3) : bid;
    if (b >= ntile) return;
    const int ti = b % tiles_i, t2 = b / tiles_i;
    const int tj = t2 % tiles_j, tk = t2 / tiles_j;
    const int i0 = ti * GX, j0 = tj * GY;
    const int i = i0 + tx, j = j0 + ty;
    const int kb = tk * kc, ke = min(kb + kc, a.nk) - 1;
    const bool inb = i < a.ni && j < a.nj;
    const long col = (long)min(i, a.ni - 1) + a.sj * min(j, a.nj - 1);
    auto kcl = [&](int kk) { return a.sk * (long)min(max(kk, 0), a.nk - 1); };

    if (!(a.it < *a.iters)) {
        // an iteration the forward did not run: the adjoint passes through
        if (!inb) return;
        for (int k = kb; k <= ke; ++k) {
            const long m = col + a.sk * (long)k;
            const double lam = a.lam[m];
            const double lin = 0.0;
            const double adn0 = a.stage != 3 ? a.adn[m] : 0.0;
            if (a.stage == 3) { a.adn[m] = lin; a.out[m] = lam; }
            else if (a.stage == 2) { a.adn[m] = adn0 + lin; a.out[m] = lam; }
            else a.out[m] = lam + lin + adn0;
        }
        return;
    }

    // this thread's halo cell: hkind 1 j halo, 2 i halo, 0 none
    int hkind = 0, hoff = 0, hrow = 0, hcol = 0;
    if (tid < NJH) {
        const int r = tid / GX, c = tid % GX;
        const int jj = r < GH ? j0 - GH + r : j0 + GY + (r - GH);
        hkind = 1; hrow = r < GH ? r : GY + r; hcol = c;
        hoff = min(i0 + c, a.ni - 1) + (int)a.sj * min(max(jj, 0), a.nj - 1);
    } else if (tid < NJH + NIH) {
        const int t = tid - NJH, r = t / (2 * GH), c = t % (2 * GH);
        const int ii = c < GH ? i0 - GH + c : i0 + GX + (c - GH);
        hkind = 2; hrow = r; hcol = c < GH ? c : GX + c;
        hoff = min(max(ii, 0), a.ni - 1) + (int)a.sj * min(j0 + r, a.nj - 1);
    }
    const double *__restrict__ hq = hkind == 2 ? a.q[0] : a.q[1];

    // operands of plane t, loaded one plane ahead: the column's lam, q, code
    // (q_0, q_1 and the halo cell only on the tile's own planes), and the
    // output-time operands (conv, adn) of plane t - 3
    double nl = 0.0, nq0 = 0.0, nq1 = 0.0, nq2 = 0.0, nhl = 0.0, nhq = 0.0, nadn = 0.0;
    int nc = 0, nhc = 0, ncv = 0;
    auto fetch = [&](int t) {
        const long l = col + kcl(t);
        nl = a.lam[l]; nq2 = a.q[2][l]; nc = a.code[l];
        if (t >= kb && t <= ke) {
            nq0 = a.q[0][l]; nq1 = a.q[1][l];
            if (hkind) {
                const long h = hoff + a.sk * (long)t;
                nhl = a.lam[h]; nhq = hq[h]; nhc = a.code[h];
            }
        }
        const int m = t - 3;
        if (m >= kb && m <= ke) {
            const long lm = col + a.sk * (long)m;
            ncv = __builtin_nontemporal_load(a.conv + lm);
            if (a.stage != 3) nadn = __builtin_nontemporal_load(a.adn + lm);
        }
    };

    // rings over the planes: lam q_2 and codes of planes t-6 .. t, lam and the
    // i/j sums of planes t-3 .. t
    double kq[7], lr[4], sr[4];
    int kcd[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) { kq[q] = 0.0; kcd[q] = 0; }
#pragma unroll
    for (int q = 0; q < 4; ++q) { lr[q] = 0.0; sr[q] = 0.0; }

    const bool j_on = j == 0 || j == a.nj - 1;
    // waves whose row or lanes reach within 3 nodes of a face take the face terms
    const bool near_i = i0 <= 3 || i0 + GX - 1 >= a.ni - 4;
    const bool near_j = j <= 3 || j >= a.nj - 4;
    fetch(kb - 3);
    for (int t = kb - 3; t <= ke + 3; ++t) {
        const double lam = nl, q0 = nq0, q1 = nq1, q2 = nq2, hl = nhl, hqv = nhq;
        const int code = nc, hc = nhc, cv = ncv;
        const double adn0 = nadn;
        if (t < ke + 3) fetch(t + 1);
        // shift the rings
#pragma unroll
        for (int q = 0; q < 6; ++q) { kq[q] = kq[q + 1]; kcd[q] = kcd[q + 1]; }
#pragma unroll
        for (int q = 0; q < 3; ++q) { lr[q] = lr[q + 1]; sr[q] = sr[q + 1]; }
        kq[6] = lam * q2; kcd[6] = code; lr[3] = lam; sr[3] = 0.0;
        if (t >= kb && t <= ke) {
            const int buf = t & 1;
            sW0[buf][ty][tx + GH] = lam * q0;
            sCi[buf][ty][tx + GH] = (uint16_t)code;
            sW1[buf][ty + GH][tx] = lam * q1;
            sCj[buf][ty + GH][tx] = (uint16_t)code;
            if (hkind == 1) { sW1[buf][hrow][hcol] = hl * hqv; sCj[buf][hrow][hcol] = (uint16_t)hc; }
            else if (hkind == 2) { sW0[buf][hrow][hcol] = hl * hqv; sCi[buf][hrow][hcol] = (uint16_t)hc; }
            __syncthreads();
            // i and j terms of this plane: the nodes whose stencil holds (i, j, t)
            double s = 0.0;
#pragma unroll
            for (int o = -2; o <= 2; ++o) {
                const int ia = i - o;
                const int c = sCi[buf][ty][tx + GH - o];
                const double e = s_eno[c & 15][o + 2];
                const bool ok = (c & kLive) && ia >= 1 && ia <= a.ni - 2;
                s += ok ? sW0[buf][ty][tx + GH - o] * e : 0.0;
            }
            if (near_i) {
                // one-sided differences of the face nodes i = 0 (slots 0..3 at
                // i = 0..3) and i = ni-1 (slots 0..3 at i = ni-1 .. ni-4)
                if (i <= 3) {
                    const int c = sCi[buf][ty][GH - i0];
                    if (c & kLive)
                        s += sW0[buf][ty][GH - i0] *
                             face_coef(c & 15, a.order_bc, 1, a.eno_bc != 0, i, t == 0 && !j_on);
                }
                if (i >= a.ni - 4 && i <= a.ni - 1) {
                    const int c = sCi[buf][ty][GH + a.ni - 1 - i0];
                    if (c & kLive)
                        s += sW0[buf][ty][GH + a.ni - 1 - i0] *
                             face_coef(c & 15, a.order_bc, -1, a.eno_bc != 0, a.ni - 1 - i, false);
                }
            }
#pragma unroll
            for (int o = -2; o <= 2; ++o) {
                const int ja = j - o;
                const int c = sCj[buf][ty + GH - o][tx];
                const double e = s_eno[(c >> 4) & 15][o + 2];
                const bool ok = (c & kLive) && ja >= 1 && ja <= a.nj - 2;
                s += ok ? sW1[buf][ty + GH - o][tx] * e : 0.0;
            }
            if (near_j) {
                if (j <= 3) {
                    const int c = sCj[buf][GH - j0][tx];
                    if (c & kLive)
                        s += sW1[buf][GH - j0][tx] * face_coef((c >> 4) & 15, a.order_bc, 1, a.eno_bc != 0, j, false);
                }
                if (j >= a.nj - 4) {
                    const int c = sCj[buf][GH + a.nj - 1 - j0][tx];
                    if (c & kLive)
                        s += sW1[buf][GH + a.nj - 1 - j0][tx] *
                             face_coef((c >> 4) & 15, a.order_bc, -1, a.eno_bc != 0, a.nj - 1 - j, false);
                }
            }
            sr[3] = s;
        }
        // plane m = t - 3 has all its k neighbours (planes m-3 .. m+3) in the ring
        const int m = t - 3;
        if (m >= kb && m <= ke && inb) {
            double s = 0.0;
#pragma unroll
            for (int o = -2; o <= 2; ++o) {
                const int ka = m - o;
                const int c = kcd[3 - o];
                const double e = s_eno[(c >> 8) & 15][o + 2];
                const bool ok = (c & kLive) && ka >= 1 && ka <= a.nk - 2;
                s += ok ? kq[3 - o] * e : 0.0;
            }
            // the face planes 0 (ring slot 3 - m) and nk-1 (ring slot 3 + nk-1-m),
            // picked by unrolled selects so the rings stay in registers
            if (m <= 3 || m >= a.nk - 4) {
                double wlo = 0.0, whi = 0.0;
                int clo = 0, chi = 0;
#pragma unroll
                for (int r = 0; r < 7; ++r) {
                    if (r == 3 - m) { wlo = kq[r]; clo = kcd[r]; }
                    if (r == 3 + (a.nk - 1 - m)) { whi = kq[r]; chi = kcd[r]; }
                }
                if (m <= 3 && (clo & kLive))
                    s += wlo * face_coef((clo >> 8) & 15, a.order_bc, 1, a.eno_bc != 0, m, false);
                if (m >= a.nk - 4 && (chi & kLive))
                    s += whi * face_coef((chi >> 8) & 15, a.order_bc, -1, a.eno_bc != 0, a.nk - 1 - m, false);
            }
            const long lm = col + a.sk * (long)m;
            const double lm_ = lr[0];
            const bool upd = cv != 0;
            double acc = upd ? a.c2 * lm_ : lm_;
            acc += sr[0] + s;
            if (kcd[3] & kHasT) a.adj_s[lm] += lm_ * a.T[lm];
            const double lin = upd ? a.c1 * lm_ : 0.0;
            if (a.stage == 3) { a.adn[lm] = lin; a.out[lm] = acc; }
            else if (a.stage == 2) { a.adn[lm] = adn0 + lin; a.out[lm] = acc; }
            else a.out[lm] = acc + lin + adn0;
        }
    }
}

// adj_phi0 = lam + adj_s * s'(phi0)
struct FinalArgs {
    long n;
    const double *lam, *adj_s, *phi0;
    double eps;
    double *adj_phi0;
};

struct KFinal {
    using Args = FinalArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long l = (long)bid * 256 + threadIdx.x;
        if (l >= a.n) return;
        a.adj_phi0[l] = a.lam[l] + a.adj_s[l] * sign_phi0_deriv(a.phi0[l], a.eps);
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void radj_final(FinalArgs a) { KFinal::run(a, (int)blockIdx.x, (int)gridDim.x); }

void KFinal::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { radj_final<<<g, b, 0, s>>>(a); }

struct ZeroArgs {
    long n;
    double *p;
};

struct KZero {
    using Args = ZeroArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long l = (long)bid * 256 + threadIdx.x;
        if (l < a.n) a.p[l] = 0.0;
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void radj_zero(long n, double *p) { KZero::run(ZeroArgs{n, p}, (int)blockIdx.x, (int)gridDim.x); }

void KZero::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { radj_zero<<<g, b, 0, s>>>(a.n, a.p); }

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }
size_t fbytes(const Blk &b) { return ((size_t)b.n * sizeof(double) + 255) & ~(size_t)255; }

}  // namespace

size_t reinit_eno2_adjoint_work(const Blk &b, int itermax)
{
    // the tape, the forward result (when the caller does not take it) and four adjoint fields
    return reinit_trace_work(b, itermax) + 5 * fbytes(b) + 4096;
}

int launch_reinit_eno2_adjoint(const Blk &b, const LsmReinitParams &p, const double *phi0,
                               const double *adj_out, double *adj_phi0, double *out, Carver &w,
                               hipStream_t s, const ReinitTape *pre)
{
    ReinitTape own;
    int rc;
    if (!pre) {
        double *res = out ? out : (double *)w.take((size_t)b.n * sizeof(double));
        if (!res) return fail(LSM_ERR_WORKSPACE, "reinit adjoint workspace too small");
        if ((rc = launch_reinit_eno2_trace(b, p, phi0, own, res, w, s))) return rc;
        pre = &own;
    }
    const ReinitTape &t = *pre;
    const size_t fb = (size_t)b.n * sizeof(double);
    double *L0 = (double *)w.take(fb), *L1 = (double *)w.take(fb);
    double *adn = (double *)w.take(fb), *adj_s = (double *)w.take(fb);
    if (!L0 || !L1 || !adn || !adj_s) return fail(LSM_ERR_WORKSPACE, "reinit adjoint workspace too small");
    if ((rc = klaunch<KZero>(nblk(b.n), dim3(256), s, ZeroArgs{b.n, adj_s}, "radj_zero"))) return rc;
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    // tile rows (8 or 16) and planes per tile of the gather
    // (planes per tile: up to 32 while the grid keeps about four workgroups
    // per CU; the k window's 7 halo planes are then read once per 32)
    int gy = 8;
    int kc = (int)std::min<long>(32, std::max<long>(8, (long)b.e.nk * ((b.e.ni + GX - 1) / GX) *
                                                            ((b.e.nj + 7) / 8) / 1024));
    if (const char *e = getenv("LSM_RADJ_GY")) gy = atoi(e) == 16 ? 16 : 8;
    if (const char *e = getenv("LSM_RADJ_KC")) { const int v = atoi(e); if (v > 0 && v <= 64) kc = v; }
    // the first stage reads adj_out itself; then the two buffers alternate
    const double *lam = adj_out;
    double *bufs[2] = {L0, L1};
    int wi = 0;
    for (int it = t.nit - 1; it >= 0; --it) {
        for (int st = 2; st >= 0; --st) {
            KTimer kt(LSM_KT_REINIT_ADJOINT, s);
            double *other = bufs[wi];
            TGatherArgs ga{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, b.n, lam,
                           {t.q[it][st][0], t.q[it][st][1], t.q[it][st][2]}, t.code[it][st], t.T[it][st],
                           t.conv[it], t.iters, it, st + 1, p.order_bc, p.eno_bc, C[st][0], C[st][1], other, adn,
                           adj_s};
            const int tiles_i = (b.e.ni + GX - 1) / GX, tiles_j = (b.e.nj + gy - 1) / gy;
            const int tiles_k = (b.e.nk + kc - 1) / kc;
            const unsigned grid = (unsigned)((tiles_i * tiles_j * tiles_k + 7) / 8 * 8);
            const TGatherLaunch gl{ga, tiles_i, tiles_j, kc};
            if (gy == 16) rc = klaunch<KTGather<16>>(grid, dim3(GX, 16, 1), s, gl, "radj_tape_gather");
            else rc = klaunch<KTGather<8>>(grid, dim3(GX, 8, 1), s, gl, "radj_tape_gather");
            if (rc) return rc;
            lam = other;
            wi ^= 1;
        }
    }
    return klaunch<KFinal>(nblk(b.n), dim3(256), s, FinalArgs{b.n, lam, adj_s, phi0, p.epslsm, adj_phi0},
                           "radj_final");
}

}  // namespace lsm
