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
__device__ __forceinline__ void radj_tape_gather_run(const TGatherArgs &a, int tiles_i, int tiles_j, int kc, int bid);

template <int GY>
struct KTGather {
    using Args = TGatherLaunch;
    static constexpr int kThreads = GX * GY;
    __device__ __forceinline__ static void run(const Args &x, int bid, int)
    {
        radj_tape_gather_run<GY>(x.a, x.tiles_i, x.tiles_j, x.kc, bid);
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
__device__ __forceinline__ void radj_tape_gather_run(const TGatherArgs &a, int tiles_i, int tiles_j, int kc, int bid)
{
    __shared__ double sW0[GY][GX + 2 * GH];
    __shared__ double sW1[GY + 2 * GH][GX];
    __shared__ uint16_t sC[GY + 2 * GH][GX + 2 * GH];
    __shared__ double s_eno[16][5];   // interior coefficients of every ENO2 branch code
    // k window of each thread's own column, planes k-3 .. k+3 in a ring of 7
    // (slot kk mod 7; only the thread itself reads its entries, and plane k+4
    // replaces plane k-3 after the plane's sums: no barrier needed)
    __shared__ double sK[7][GY][GX];
    __shared__ uint16_t sKc[7][GY][GX];
    auto rs = [](int kk) { return (kk + 70) % 7; };
    constexpr int NT = GX * GY;
    constexpr int N0 = GY * (GX + 2 * GH), N1 = (GY + 2 * GH) * GX, NC = (GY + 2 * GH) * (GX + 2 * GH);
    constexpr int R0 = (N0 + NT - 1) / NT, R1 = (N1 + NT - 1) / NT, RC = (NC + NT - 1) / NT;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * GX + tx;
    if (tid < 16) {
        double e[5];
        eno2_coef(tid, e);
        for (int q = 0; q < 5; ++q) s_eno[tid][q] = e[q];
    }
    const int b = bid;
    const int ti = b % tiles_i, t2 = b / tiles_i;
    const int tj = t2 % tiles_j, tk = t2 / tiles_j;
    const int i0 = ti * GX, j0 = tj * GY;
    const int i = i0 + tx, j = j0 + ty;
    const int kb = tk * kc, ke = min(kb + kc, a.nk) - 1;
    const bool inb = i < a.ni && j < a.nj;
    const int ic = min(i, a.ni - 1), jc = min(j, a.nj - 1);
    const long col = (long)ic + a.sj * jc;
    const bool active = a.it < *a.iters;
    const bool j_on = j == 0 || j == a.nj - 1;
    auto kcl = [&](int kk) { return min(max(kk, 0), a.nk - 1); };
    // in-plane offsets of this thread's halo-fill entries (the same on every plane)
    int o0[R0], o1[R1], oc[RC];
#pragma unroll
    for (int r = 0; r < R0; ++r) {
        const int c = min(tid + r * NT, N0 - 1), rr = c / (GX + 2 * GH), cc = c % (GX + 2 * GH);
        o0[r] = min(max(i0 - GH + cc, 0), a.ni - 1) + (int)a.sj * min(j0 + rr, a.nj - 1);
    }
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        const int c = min(tid + r * NT, N1 - 1), rr = c / GX, cc = c % GX;
        o1[r] = min(i0 + cc, a.ni - 1) + (int)a.sj * min(max(j0 - GH + rr, 0), a.nj - 1);
    }
#pragma unroll
    for (int r = 0; r < RC; ++r) {
        const int c = min(tid + r * NT, NC - 1), rr = c / (GX + 2 * GH), cc = c % (GX + 2 * GH);
        oc[r] = min(max(i0 - GH + cc, 0), a.ni - 1) + (int)a.sj * min(max(j0 - GH + rr, 0), a.nj - 1);
    }
    // k window: planes k-3 .. k+3 of this column (clamped; entries outside the
    // block are never used)
    for (int q = 0; q < 7; ++q) {
        const int kk = kb - 3 + q;
        const long l = col + a.sk * (long)kcl(kk);
        sK[rs(kk)][ty][tx] = a.lam[l] * a.q[2][l];
        sKc[rs(kk)][ty][tx] = a.code[l];
    }
    // the halo planes of the next plane are loaded into registers while the
    // current plane's sums run, and stored to LDS after the plane's barrier
    double l0[R0], q0[R0], l1[R1], q1[R1];
    uint16_t cv[RC];
    auto fetch = [&](int k) {
        const double *lk = a.lam + a.sk * (long)k, *q0k = a.q[0] + a.sk * (long)k, *q1k = a.q[1] + a.sk * (long)k;
        const uint16_t *ck = a.code + a.sk * (long)k;
#pragma unroll
        for (int r = 0; r < R0; ++r) { l0[r] = lk[o0[r]]; q0[r] = q0k[o0[r]]; }
#pragma unroll
        for (int r = 0; r < R1; ++r) { l1[r] = lk[o1[r]]; q1[r] = q1k[o1[r]]; }
#pragma unroll
        for (int r = 0; r < RC; ++r) cv[r] = ck[oc[r]];
    };
    fetch(kb);
    for (int k = kb; k <= ke; ++k) {
        const long pk = a.sk * (long)k;
        double *w0 = &sW0[0][0], *w1 = &sW1[0][0];
        uint16_t *wc = &sC[0][0];
#pragma unroll
        for (int r = 0; r < R0; ++r)
            if (tid + r * NT < N0) w0[tid + r * NT] = l0[r] * q0[r];
#pragma unroll
        for (int r = 0; r < R1; ++r)
            if (tid + r * NT < N1) w1[tid + r * NT] = l1[r] * q1[r];
#pragma unroll
        for (int r = 0; r < RC; ++r)
            if (tid + r * NT < NC) wc[tid + r * NT] = cv[r];
        __syncthreads();
        const long m = col + pk;
        const double lam = a.lam[m];
        const uint8_t cnv = a.conv[m];
        double adn0 = 0.0;
        if (a.stage != 3) adn0 = a.adn[m];
        // the next plane of the k window and of the halos
        const long lk = col + a.sk * (long)kcl(k + 4);
        const double nlk = a.lam[lk], nqk = a.q[2][lk];
        const uint16_t nck = a.code[lk];
        if (k < ke) fetch(k + 1);
        if (inb) {
            const bool upd = active && cnv;
            double acc = upd ? a.c2 * lam : lam;
            if (active) {
                if (sC[ty + GH][tx + GH] & kHasT) a.adj_s[m] += lam * a.T[m];
                const int idx[3] = {i, j, k};
                const int nn[3] = {a.ni, a.nj, a.nk};
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
                        int c;
                        double w;
                        if (d == 0) { c = sC[ty + GH][tx + GH - o]; w = sW0[ty][tx + GH - o]; }
                        else if (d == 1) { c = sC[ty + GH - o][tx + GH]; w = sW1[ty + GH - o][tx]; }
                        else { c = sKc[rs(k - o)][ty][tx]; w = sK[rs(k - o)][ty][tx]; }
                        if (!(c & kLive)) continue;
                        const int br = (c >> (4 * d)) & 15;
                        const double e = kind == 0 ? s_eno[br][slot]
                                                   : face_coef(br, a.order_bc, kind == 1 ? 1 : -1, a.eno_bc != 0, slot,
                                                               d == 0 && ia == 0 && k == 0 && !j_on);   // edge i = 1, k = 1
                        acc += w * e;
                    }
                }
            }
            const double lin = upd ? a.c1 * lam : 0.0;
            if (a.stage == 3) { a.adn[m] = lin; a.out[m] = acc; }
            else if (a.stage == 2) { a.adn[m] = adn0 + lin; a.out[m] = acc; }
            else a.out[m] = acc + lin + adn0;
        }
        // slide the k window: plane k + 4 replaces plane k - 3
        sK[rs(k + 4)][ty][tx] = nlk * nqk;
        sKc[rs(k + 4)][ty][tx] = nck;
        __syncthreads();
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
            const unsigned grid = (unsigned)(tiles_i * tiles_j * tiles_k);
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
