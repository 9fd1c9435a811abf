// Batched multi-block level-set advection: many curvilinear blocks, of uneven
// sizes, advanced together on one GPU.
//
// The blocks are either the blocks of one multiblock domain (a lattice of
// blocks with matching face sizes, `ghost` layers on every face, the MPI
// block decomposition of the reference, init.F90:489-1068, held on one
// device) or independent instances (no ghost layers, every face physical).
// One RK stage (ad_RKTVD3.F90:62-84) is three launches:
//   * the interior tiles of every block (64 x TY columns x kc planes, the
//     single-block stage tile), pulled by a persistent grid from eight work
//     queues, one per XCD (workgroup b runs on XCD b % 8 and starts on queue
//     b % 8, so consecutive tiles of one block share that XCD's L2), with
//     work stealing from the other queues once its own is empty; or the same
//     tiles as one workgroup each (`grid` form);
//   * the physical-face shell of every block (RK update with rh = 0 and
//     bcond_lsm, bcond_lsm.F90:26-287);
//   * the halo fill (rhs_exchng3_3d.F90:6-359): every ghost node that lies
//     inside the domain takes the value of the block that owns that node,
//     which is what the reference's i, j, k exchanges of full ghost slabs
//     produce, edges and corners included.
// The per-block form launches the single-block kernels block by block (the
// reference's one-rank-per-block structure without MPI), for comparison.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "lsm_advect_tile.hpp"
#include "lsm_internal.hpp"
#include "lsm_reinit_common.hpp"

namespace lsm {

namespace {

constexpr int kMaxLattice = 32;   // blocks per lattice axis

struct BBlock {
    long off;                 // element offset of the block's array in the batch arrays
    int ni, nj, nk;           // array extents incl. ghost layers
    long sj, sk;
    Range r;                  // owned / right-hand-side nodes, physical faces (local indices)
    int gs[3];                // global index of the owned node (ia, ja, ka)
    int bt[6];
    double big, dc, de, dz;
    int tiles_i, tiles_j, ntiles, tile0, kc, jb0, ib0;
    long shell0;              // first shell node of the block in the batch enumeration
    long nI, nJ, nK;
    int cnt_i, cnt_j;
    int rt_tiles_i, rt_tiles_j, rt_ntiles, rt_tile0, rt_kc;   // reinitialisation tiles (W x 256/W x rt_kc)
    int rt_ib0;               // i of the first reinitialisation tile column (on a 128-byte line)
    int sh_wg0;               // first shell workgroup of the block (reinitialisation)
};

struct BFields {
    const double *ps, *pn, *U1, *U2, *U3, *aj;
    double *out;
    double c1, c2, c3, dt;
};

__device__ __forceinline__ StageArgs stage_args(const BBlock &B, const BFields &f)
{
    StageArgs a;
    a.ps = f.ps + B.off; a.pn = f.pn + B.off; a.out = f.out + B.off;
    a.U1 = f.U1 + B.off; a.U2 = f.U2 + B.off; a.U3 = f.U3 + B.off; a.aj = f.aj + B.off;
    a.mask = nullptr;
    a.ni = B.ni; a.nj = B.nj; a.nk = B.nk; a.sj = B.sj; a.sk = B.sk;
    a.r = B.r;
    a.c1 = f.c1; a.c2 = f.c2; a.c3 = f.c3; a.dt = f.dt; a.dc = B.dc; a.de = B.de; a.dz = B.dz;
    a.tiles_i = B.tiles_i; a.tiles_j = B.tiles_j; a.kc = B.kc; a.ntiles = B.ntiles; a.jb0 = B.jb0; a.ib0 = B.ib0;
    return a;
}

struct QueueArgs {
    const BBlock *blocks;
    const int32_t *tile_block;   // block of every tile
    int *qcount;                 // 8 counters, zeroed before the launch
    int qstart[9];               // tiles of queue q: qstart[q] .. qstart[q+1]-1
    BFields f;
};

// persistent: each workgroup pops tiles until every queue is empty
template <bool FAST, int TY, int W>
__global__ __launch_bounds__(TY * W) void batch_stage_queue(QueueArgs a)
{
    __shared__ int s_tile;
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    int q = blockIdx.x & 7, tried = 0;
    for (;;) {
        if (lead) {
            int t = -1;
            while (tried < 8) {
                const int qq = (q + tried) & 7;
                const int len = a.qstart[qq + 1] - a.qstart[qq];
                if (len > 0) {
                    const int k = atomicAdd(&a.qcount[qq], 1);
                    if (k < len) { t = a.qstart[qq] + k; break; }
                }
                ++tried;
            }
            s_tile = t;
        }
        __syncthreads();
        const int t = s_tile;
        __syncthreads();
        if (t < 0) break;
        const BBlock &B = a.blocks[a.tile_block[t]];
        advect_tile<FAST, false, TY, true, false, W>(stage_args(B, a.f), t - B.tile0);
        __syncthreads();   // the next tile reuses the LDS planes
    }
}

// one workgroup per tile
template <bool FAST, int TY, int W>
__global__ __launch_bounds__(TY * W) void batch_stage_grid(QueueArgs a, int ntiles)
{
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= ntiles) return;
    const BBlock &B = a.blocks[a.tile_block[t]];
    advect_tile<FAST, false, TY, true, false, W>(stage_args(B, a.f), t - B.tile0);
}

__global__ void batch_queue_reset(int *qcount)
{
    if (threadIdx.x < 8) qcount[threadIdx.x] = 0;
}

// block whose shell-node range [shell0, next shell0) holds t
__device__ __forceinline__ int find_block(const BBlock *blocks, int nb, long t)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long s = blocks[mid].shell0;
        if (s <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct ShellBArgs {
    const BBlock *blocks;
    int b0, nb;         // blocks b0 .. b0+nb-1
    long total;         // shell nodes of those blocks
    long base;          // shell0 of block b0
    BFields f;
};

__global__ __launch_bounds__(256) void batch_shell(ShellBArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.total) return;
    const long tg = t + a.base;
    const int b = a.b0 + find_block(a.blocks + a.b0, a.nb, tg);
    const BBlock &B = a.blocks[b];
    ShellArgs sh;
    sh.out_ro = a.f.out + B.off; sh.out = a.f.out + B.off;
    sh.ps = a.f.ps + B.off; sh.pn = a.f.pn + B.off; sh.aj = a.f.aj + B.off;
    sh.mask = nullptr;
    sh.ni = B.ni; sh.nj = B.nj; sh.nk = B.nk; sh.sj = B.sj; sh.sk = B.sk; sh.r = B.r;
    for (int q = 0; q < 6; ++q) sh.bt[q] = B.bt[q];
    sh.big = B.big; sh.c1 = a.f.c1; sh.c2 = a.f.c2; sh.c3 = a.f.c3; sh.dt = a.f.dt;
    sh.nI = B.nI; sh.nJ = B.nJ; sh.nK = B.nK; sh.cnt_i = B.cnt_i; sh.cnt_j = B.cnt_j;
    int i, j, k;
    if (!shell_node(sh, tg - B.shell0, i, j, k)) return;
    const double v = sh_f3<false>(sh, i, j, k);
    sh.out[(long)i + sh.sj * j + sh.sk * k] = v;
}

// One internal-face ghost slab of one block: i faces span all (j,k) of the
// array, j faces the owned i and all k, k faces the owned (i,j).
struct HSeg {
    int b, slab;        // block, slab 0..5 = i low, i high, j low, j high, k low, k high
    int n0, n1;         // slab extents along its first two index directions (fastest first)
    int i0, j0, k0;     // local index of the slab's first node
    int count, wg0;     // nodes, first workgroup
};

struct HaloArgs {
    const BBlock *blocks;
    const HSeg *segs;
    const int32_t *wg_seg;                 // segment of every workgroup
    int wg_base;                           // workgroup of blockIdx.x = 0
    double *field;
    int g;                                 // ghost layers
    int lat[3];                            // lattice extents
    int lstart[3][kMaxLattice + 1];        // global start of each lattice slab per axis
    const int32_t *block_of;               // block id of each lattice position
};

__device__ __forceinline__ int lattice_of(const HaloArgs &a, int ax, int G)
{
    int L = 0;
    while (L + 1 < a.lat[ax] && a.lstart[ax][L + 1] <= G) ++L;
    return L;
}

__global__ __launch_bounds__(256) void batch_halo(HaloArgs a)
{
    const int w = (int)blockIdx.x + a.wg_base;
    const HSeg S = a.segs[a.wg_seg[w]];
    const int e = (w - S.wg0) * 256 + (int)threadIdx.x;
    if (e >= S.count) return;
    const BBlock &B = a.blocks[S.b];
    // node e of the slab, i fastest (n0 along i, n1 along j)
    const int q0 = e % S.n0, r = e / S.n0;
    const int i = S.i0 + q0, j = S.j0 + r % S.n1, k = S.k0 + r / S.n1;
    const int g = a.g;
    const int loc[3] = {i, j, k};
    int G[3], L[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        G[ax] = B.gs[ax] + (loc[ax] - g);
        if (G[ax] < 0 || G[ax] >= a.lstart[ax][a.lat[ax]]) return;   // outside the domain
        L[ax] = lattice_of(a, ax, G[ax]);
    }
    const int ob = a.block_of[L[0] + a.lat[0] * (L[1] + a.lat[1] * L[2])];
    const BBlock &O = a.blocks[ob];
    const long src = O.off + (long)(G[0] - a.lstart[0][L[0]] + g) + O.sj * (G[1] - a.lstart[1][L[1]] + g) +
                     O.sk * (G[2] - a.lstart[2][L[2]] + g);
    a.field[B.off + (long)i + B.sj * j + B.sk * k] = a.field[src];
}

// ---------------------------------------------------------------------------
// batched reinitialisation (reinitialisation_test2.F90:42-210 on every block
// as one rank of the reference runs it): owned nodes are updated, ghost
// layers keep their values of the start (no exchange inside the loop), |grad
// d| of the ghost nodes counts as 0 in the exit test (so a block with ghost
// layers runs every iteration), each block has its own exit test.
// ---------------------------------------------------------------------------
struct RCtl {
    unsigned long long maxbits, minbits;
    int active, iters;
};

constexpr unsigned long long kInf64 = 0x7ff0000000000000ull;

struct RStage {
    double c1, c2, c3, dt, eps, eps_r;
    int order_bc, eno_bc;
};

struct BRArgs {
    const BBlock *blocks;
    const int32_t *rtile_block;   // block of every reinitialisation tile
    const int32_t *sh_block;      // block of every shell workgroup
    int tile_base, sh_base;       // first tile / shell workgroup of this launch
    int ntiles, grid_tiles;       // tiles; tile workgroups (padded to 8), shell workgroups follow
    const double *gt;             // metric tensor of every node: component c at gt + c * gs
    long gs;
    const double *din, *dn, *phi0;
    double *dout;
    uint8_t *conv;
    RCtl *ctl;                    // one per block
    RStage p;
    int reduce;                   // max / min of |grad d| for the exit test (0: the test cannot pass)
};

// |grad d| at an owned node (calc_NormGrad_d.F90:39-2446 on one rank): one-sided
// formulas on the block's physical faces, ENO2 elsewhere (ghost values in the
// stencil at internal faces); phi0s is phizero at the block's last interior
// node, which the i-face formulas use for their upwind sign.  |grad d| from
// the metric tensor gt (component c at gt + c * gs), as in the interior.
__device__ double normgrad_node_r(const double *__restrict__ d, const double *__restrict__ phi0,
                                  const double *gt, long gs, long sj, long sk, const Range &r, const double dd[3],
                                  double phi0s, int order_bc, bool eno_bc, int i, int j, int k, long l)
{
    const int idx[3] = {i, j, k};
    const long st[3] = {1, sj, sk};
    const int lo[3] = {r.ia, r.ja, r.ka}, hi[3] = {r.ib, r.jb, r.kb};
    const bool flo[3] = {r.fb, r.fl, r.fd}, fhi[3] = {r.ff, r.fr, r.fu};
    bool onl[3], onh[3], on[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        onl[a] = flo[a] && idx[a] == lo[a];
        onh[a] = fhi[a] && idx[a] == hi[a];
        on[a] = onl[a] || onh[a];
    }
    const double p0 = (on[0] && !on[1] && !on[2]) ? phi0s : phi0[l];
    double gr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long s = st[a];
        if (on[a]) {
            const int dir = onl[a] ? 1 : -1;
            const long ss = s * dir;
            if (a == 0 && onl[0] && onl[2] && !on[1]) {
                // edge i = 1, k = 1 (calc_NormGrad_d.F90:1281-1283)
                const double dc2 = 0.5 * dd[0];
                gr[a] = dc2 * ((-(1.0 * d[l + 2]) + 4.0 * d[l + 1]) - 3.0 * d[l]);
            } else {
                gr[a] = dd[a] * biased(d[l], d[l + ss], d[l + 2 * ss], d[l + 3 * ss], order_bc, dir, eno_bc);
            }
        } else {
            const bool lo_o = flo[a] && idx[a] == lo[a] + 1, hi_o = fhi[a] && idx[a] == hi[a] - 1;
            const double fLL = lo_o ? 0.0 : d[l - 2 * s];
            const double fRR = hi_o ? 0.0 : d[l + 2 * s];
            gr[a] = dd[a] * eno2(p0, fLL, d[l - s], d[l], d[l + s], fRR, lo_o, hi_o);
        }
    }
    double G[6], t[3];
#pragma unroll
    for (int c = 0; c < 6; ++c) G[c] = gt[c * gs + l];
    return rn::tensor_norm(G, gr[0], gr[1], gr[2], t);
}

// the stage update at one node (reinitialisation_test2.F90:137-181)
__device__ __forceinline__ double rn_update(const RStage &p, double ngv, double dv, double dnv, double p0, bool upd)
{
    if (!upd) return dv;
    return sussman_update(p.c1, p.c2, p.c3 * p.dt, ngv, dv, dnv, sign_phi0(p0, p.eps), deltaf(dv, p.eps));
}

__device__ __forceinline__ void rn_reduce(double vmax, double vmin, int tid, int nthreads, RCtl *ctl)
{
    for (int off = 32; off > 0; off >>= 1) {
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
        vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
    }
    __shared__ double smax[16], smin[16];
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) { smax[wv] = vmax; smin[wv] = vmin; }
    __syncthreads();
    if (tid == 0) {
        double a = smax[0], b = smin[0];
        for (int q = 1; q < nthreads / 64; ++q) { a = fmax(a, smax[q]); b = fmin(b, smin[q]); }
        atomicMax(&ctl->maxbits, (unsigned long long)__double_as_longlong(a));
        atomicMin(&ctl->minbits, (unsigned long long)__double_as_longlong(b));
    }
}

// STAGE 0: max / min of |grad din| only; 1..3: the RK stage.  Interior tiles
// of W (i) x 256/W (j) columns; W = 64 puts one row in a wavefront, W = 32
// and 16 put 2 and 4 rows in one (fewer idle lanes on blocks whose i extent
// is not a multiple of 64, but more partly used cache lines per load: on the
// 64-block workload 64 columns are fastest, 30.3 / 30.7 / 34.1 ms per full
// step for 64 / 32 / 16).
template <int STAGE, int W, bool RED>
__global__ __launch_bounds__(256) void batch_reinit_stage(BRArgs a)
{
    constexpr int TY = 256 / W;
    const int tid = threadIdx.y * W + threadIdx.x;
    const double inf = __longlong_as_double((long long)kInf64);
    double vmax = 0.0, vmin = inf;
    const int bid = blockIdx.x;
    if (bid >= a.grid_tiles) {
        // ---- owned nodes on physical faces: one node per thread ----
        const int w = bid - a.grid_tiles + a.sh_base;
        const int b = a.sh_block[w];
        const BBlock B = a.blocks[b];
        RCtl *ctl = a.ctl + b;
        if (!ctl->active) return;
        const long e = (long)(w - B.sh_wg0) * 256 + tid;
        ShellArgs sh;
        sh.r = B.r; sh.nI = B.nI; sh.nJ = B.nJ; sh.nK = B.nK; sh.cnt_i = B.cnt_i; sh.cnt_j = B.cnt_j;
        int i, j, k;
        if (shell_node(sh, e, i, j, k)) {
            const long l = (long)i + B.sj * j + B.sk * k;
            const double *din = a.din + B.off, *phi0 = a.phi0 + B.off;
            const double dd[3] = {B.dc, B.de, B.dz};
            const double p0s = phi0[(long)B.r.ihi + B.sj * B.r.jhi + B.sk * B.r.khi];
            const double ngv = normgrad_node_r(din, phi0, a.gt + B.off, a.gs, B.sj, B.sk, B.r, dd, p0s, a.p.order_bc,
                                               a.p.eno_bc != 0,
                                               i, j, k, l);
            if (STAGE > 0) {
                bool upd;
                if (STAGE == 1) { upd = ngv - 1.0 > a.p.eps_r; a.conv[B.off + l] = upd ? 1 : 0; }
                else upd = a.conv[B.off + l] != 0;
                const double dv = din[l];
                const double dnv = (STAGE == 1) ? dv : a.dn[B.off + l];
                a.dout[B.off + l] = rn_update(a.p, ngv, dv, dnv, phi0[l], upd);
            }
            vmax = ngv; vmin = ngv;
        }
        if constexpr (RED) rn_reduce(vmax, vmin, tid, 256, ctl);
        return;
    }
    // ---- interior tiles: W (i) x TY (j) columns marching up kc planes ----
    int tile = xcd_tile(bid, a.grid_tiles);
    const bool live = tile < a.ntiles;
    tile += a.tile_base;
    const int b = a.rtile_block[live ? tile : a.tile_base];
    // a copy: the descriptor then lives in scalar registers for the whole
    // tile instead of being re-read after every store (no alias analysis
    // separates it from the output arrays)
    const BBlock B = a.blocks[b];
    RCtl *ctl = a.ctl + b;
    if (!ctl->active) return;
    const int t = live ? tile - B.rt_tile0 : 0;
    const int ti = t % B.rt_tiles_i, t2 = t / B.rt_tiles_i;
    const int tj = t2 % B.rt_tiles_j, tk = t2 / B.rt_tiles_j;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const Range &r = B.r;
    const int i0 = B.rt_ib0 + ti * W;
    const int i = i0 + tx, j = r.jlo + tj * TY + ty;
    const int kb = r.klo + tk * B.rt_kc;
    const int ke = live ? min(kb + B.rt_kc - 1, r.khi) : kb - 1;
    __shared__ double sm[2][TY + 4][W + 4];
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const long sj = B.sj, sk = B.sk;
    const int ic = min(i, ni - 1), jc = min(j, nj - 1);
    const long col = (long)ic + sj * jc;
    // j halo: rows -2, -1, TY, TY+1 of the tile, loaded by the first 4 rows of threads
    const bool has_jh = ty < 4;
    const int rr = (ty < 2) ? ty : ty - 2 + TY + 2;
    const int jh = min(max(r.jlo + tj * TY - 2 + rr, 0), nj - 1);
    const long col_jh = (long)ic + sj * jh;
    // i halo: columns -2, -1, W, W+1 of every tile row, loaded by threads 0 .. 4 TY - 1
    const int lin = tid;
    const bool has_ih = lin < 4 * TY;
    const int hr = min(lin >> 2, TY - 1), hc4 = lin & 3;
    const int hcol = (hc4 < 2) ? hc4 : hc4 + W;
    const int hi_i = min(max(i0 - 2 + hcol, 0), ni - 1);
    const int hi_j = min(r.jlo + tj * TY + hr, nj - 1);
    const long col_ih = (long)hi_i + sj * hi_j;
    auto kcl = [&](int kk) -> long { return sk * (long)min(max(kk, 0), nk - 1); };
    const bool act = live && i >= r.ilo && i <= r.ihi && j >= r.jlo && j <= r.jhi;
    const bool ilo = r.fb && i == r.ilo, ihi = r.ff && i == r.ihi, jlo = r.fl && j == r.jlo, jhi = r.fr && j == r.jhi;
    const double *__restrict__ d = a.din + B.off;
    const double *phi0 = a.phi0 + B.off;
    const double dd0 = B.dc, dd1 = B.de, dd2 = B.dz;
    double q0 = d[col + kcl(kb - 2)], q1 = d[col + kcl(kb - 1)];
    double q2 = d[col + kcl(kb)], q3 = d[col + kcl(kb + 1)];
    double nq4 = d[col + kcl(kb + 2)];
    double njh = has_jh ? d[col_jh + kcl(kb)] : 0.0;
    double nih = has_ih ? d[col_ih + kcl(kb)] : 0.0;
    int buf = 0;
    for (int k = kb; k <= ke; ++k) {
        const double q4 = nq4;
        sm[buf][ty + 2][tx + 2] = q2;
        if (has_jh) sm[buf][rr][tx + 2] = njh;
        if (has_ih) sm[buf][hr + 2][hcol] = nih;
        __syncthreads();
        if (k < ke) {
            const long kn = kcl(k + 1);
            nq4 = d[col + kcl(k + 3)];
            if (has_jh) njh = d[col_jh + kn];
            if (has_ih) nih = d[col_ih + kn];
        }
        if (act) {
            const long l = col + sk * (long)k;
            const double p0 = phi0[l];
            double G[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) G[c] = __builtin_nontemporal_load(a.gt + c * a.gs + B.off + l);
            double dnv = q2;
            uint8_t cv = 0;
            if (STAGE > 1) { dnv = __builtin_nontemporal_load(a.dn + B.off + l); cv = a.conv[B.off + l]; }
            const double (*pl)[W + 4] = sm[buf];
            const int cx = tx + 2, cy = ty + 2;
            const double g0 = dd0 * eno2(p0, ilo ? 0.0 : pl[cy][cx - 2], pl[cy][cx - 1], q2, pl[cy][cx + 1],
                                         ihi ? 0.0 : pl[cy][cx + 2], ilo, ihi);
            const double g1 = dd1 * eno2(p0, jlo ? 0.0 : pl[cy - 2][cx], pl[cy - 1][cx], q2, pl[cy + 1][cx],
                                         jhi ? 0.0 : pl[cy + 2][cx], jlo, jhi);
            const bool klo = r.fd && k == r.klo, khi = r.fu && k == r.khi;
            const double g2 = dd2 * eno2(p0, klo ? 0.0 : q0, q1, q2, q3, khi ? 0.0 : q4, klo, khi);
            double t[3];
            const double ngv = rn::tensor_norm(G, g0, g1, g2, t);
            if (STAGE > 0) {
                bool upd;
                if (STAGE == 1) { upd = ngv - 1.0 > a.p.eps_r; a.conv[B.off + l] = upd ? 1 : 0; }
                else upd = cv != 0;
                __builtin_nontemporal_store(rn_update(a.p, ngv, q2, dnv, p0, upd), a.dout + B.off + l);
            }
            if constexpr (RED) { vmax = fmax(vmax, ngv); vmin = fmin(vmin, ngv); }
        }
        q0 = q1; q1 = q2; q2 = q3; q3 = q4;
        buf ^= 1;
    }
    if constexpr (RED) rn_reduce(vmax, vmin, tid, 256, ctl);
}

// per-block exit-test state: a block with ghost layers counts |grad d| = 0
// there, so its minimum starts at 0
__global__ void batch_reinit_init(RCtl *ctl, int nb, int ghosted, int32_t *iters)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    ctl[b].maxbits = 0ull;
    ctl[b].minbits = ghosted ? 0ull : kInf64;
    ctl[b].active = 1;
    ctl[b].iters = 0;
    if (iters) iters[b] = 0;
}

// exit test at the top of a pseudo-time iteration (reinitialisation_test2.F90:100-107), per block
__global__ void batch_reinit_check(RCtl *ctl, int nb, int ghosted, double eps, int32_t *iters)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    RCtl &c = ctl[b];
    if (c.active) {
        const double mx = __longlong_as_double((long long)c.maxbits);
        const double mn = __longlong_as_double((long long)c.minbits);
        if (fabs(mx - 1.0) < eps && fabs(mn - 1.0) < eps) c.active = 0;
        else c.iters += 1;
    }
    c.maxbits = 0ull;
    c.minbits = ghosted ? 0ull : kInf64;
    if (iters) iters[b] = c.iters;
}

// ghost layers of src copied into dst (same block, same positions)
__global__ __launch_bounds__(256) void batch_ghost_copy(HaloArgs a, const double *src)
{
    const int w = (int)blockIdx.x + a.wg_base;
    const HSeg S = a.segs[a.wg_seg[w]];
    const int e = (w - S.wg0) * 256 + (int)threadIdx.x;
    if (e >= S.count) return;
    const BBlock &B = a.blocks[S.b];
    const int q0 = e % S.n0, r = e / S.n0;
    const long l = B.off + (long)(S.i0 + q0) + B.sj * (S.j0 + r % S.n1) + B.sk * (S.k0 + r / S.n1);
    a.field[l] = src[l];
}

// bcond_lsm.F90:26-287 on the physical faces of one axis for every block, in
// place: faces of one axis read only nodes no face of that axis writes
struct BcSegArgs {
    const BBlock *blocks;
    const HSeg *segs;            // slab 0/1 (axis low/high face), n0 x n1 owned nodes of the other two axes
    const int32_t *wg_seg;
    int wg_base, axis;
    double *field;
};

__global__ __launch_bounds__(256) void batch_bcond_phase(BcSegArgs a)
{
    const int w = (int)blockIdx.x + a.wg_base;
    const HSeg S = a.segs[a.wg_seg[w]];
    const int e = (w - S.wg0) * 256 + (int)threadIdx.x;
    if (e >= S.count) return;
    const BBlock &B = a.blocks[S.b];
    const int q0 = e % S.n0, q1 = e / S.n0;
    const int ax = a.axis;
    int i, j, k;
    long s;
    const bool hi = S.slab & 1;
    if (ax == 0) { i = hi ? B.r.ib : B.r.ia; j = B.r.ja + q0; k = B.r.ka + q1; s = 1; }
    else if (ax == 1) { j = hi ? B.r.jb : B.r.ja; i = B.r.ia + q0; k = B.r.ka + q1; s = B.sj; }
    else { k = hi ? B.r.kb : B.r.ka; i = B.r.ia + q0; j = B.r.ja + q1; s = B.sk; }
    const int bt = B.bt[2 * ax + (hi ? 1 : 0)];
    if (bt == 2 && !(ax == 0 && hi)) return;
    double *f = a.field + B.off;
    const long l = (long)i + B.sj * j + B.sk * k;
    const long d = hi ? -s : s;
    f[l] = bc_face(bt, f[l + d], f[l + 2 * d], f[l], B.big, ax == 0 && hi);
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct BatchTune {
    int form;        // LSM_BATCH_QUEUE, LSM_BATCH_GRID, LSM_BATCH_PERBLOCK
    int w;           // tile columns (64 or 32)
    int ty;          // tile rows (4 or 8)
    int kc;          // planes per tile
    int wg_per_cu;   // persistent workgroups per CU (queue form)
};

}  // namespace lsm

struct LsmBatch {
    int dev;
    int nb, g, connected;
    int lat[3];
    int lstart[3][lsm::kMaxLattice + 1];
    long total;
    std::vector<LsmBatchBlock> desc;
    std::vector<lsm::BBlock> hb;
    lsm::BBlock *db = nullptr;
    int32_t *d_tile_block = nullptr;
    int32_t *d_block_of = nullptr;
    const double *met[9] = {nullptr};  // batch metric arrays (csi, eta, zet), may be NULL
    double *gten = nullptr;            // metric tensor of every node (6 x total), with the metrics
    int32_t *d_rtile_block = nullptr;  // reinitialisation tiles
    int rt_ntiles = 0;
    int rt_w = 64;                     // reinitialisation tile columns (64, 32 or 16; rows 256 / rt_w)
    int32_t *d_sh_block = nullptr;     // shell workgroups (reinitialisation)
    int sh_wgs = 0;
    lsm::HSeg *d_bseg[3] = {nullptr};  // bcond_lsm face segments per axis
    int32_t *d_bwg[3] = {nullptr};
    std::vector<int> bc_wg[3];         // first bcond workgroup of each block per axis (nb + 1)
    lsm::HSeg *d_segs = nullptr;       // halo segments of every block
    int32_t *d_wg_seg = nullptr;       // segment of every halo workgroup
    int halo_wgs = 0;
    std::vector<int> blk_wg;           // first halo workgroup of each block (nb + 1 entries)
    int *d_qcount = nullptr;
    int qstart[9];
    int ntiles = 0;
    long nshell = 0;
    const double *aj = nullptr;
    lsm::BatchTune tune{LSM_BATCH_QUEUE, 64, 4, 8, 0};
    int cus = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace lsm {

namespace {

Range block_range(const LsmBatchBlock &d, int g)
{
    Range r;
    const int f = d.phys_faces;
    r.fb = f & LSM_FACE_I1; r.ff = f & LSM_FACE_IM;
    r.fl = f & LSM_FACE_J1; r.fr = f & LSM_FACE_JM;
    r.fd = f & LSM_FACE_K1; r.fu = f & LSM_FACE_KM;
    r.ia = g; r.ib = g + d.ni - 1;
    r.ja = g; r.jb = g + d.nj - 1;
    r.ka = g; r.kb = g + d.nk - 1;
    r.ilo = r.ia + (r.fb ? 1 : 0); r.ihi = r.ib - (r.ff ? 1 : 0);
    r.jlo = r.ja + (r.fl ? 1 : 0); r.jhi = r.jb - (r.fr ? 1 : 0);
    r.klo = r.ka + (r.fd ? 1 : 0); r.khi = r.kb - (r.fu ? 1 : 0);
    return r;
}

// A captured step holds the device table pointers, grid sizes and queue
// bounds of the configuration it was recorded with: any reconfiguration
// drops it (lsm_batch_graph_launch then fails until the next capture).  The
// caller has synchronised the device, so no replay is in flight.
void drop_graph(LsmBatch *B)
{
    if (B->gexec) { (void)hipGraphExecDestroy(B->gexec); B->gexec = nullptr; }
    if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
}

// (re)builds the tile tables for tiles of w columns, ty rows and kc planes
int build_tiles(LsmBatch *B, int w, int ty, int kc)
{
    drop_graph(B);
    std::vector<int32_t> tb;
    for (int b = 0; b < B->nb; ++b) {
        BBlock &X = B->hb[b];
        const Range &r = X.r;
        // tiles start at the last column at or before the first owned one
        // that lies on a 128-byte line (in the first row; the padded layouts
        // align every row): ghost columns left of it are never computed
        X.ib0 = max(0, r.ia - (int)((X.off + r.ia) % 16));
        X.tiles_i = (r.ihi - X.ib0 + 1 + w - 1) / w;
        X.jb0 = r.jlo;
        const int nj_int = r.jhi - r.jlo + 1, nk_int = r.khi - r.klo + 1;
        X.tiles_j = nj_int > 0 ? (nj_int + ty - 1) / ty : 0;
        X.kc = kc;
        const int tiles_k = nk_int > 0 ? (nk_int + kc - 1) / kc : 0;
        X.ntiles = (r.ihi >= r.ilo) ? X.tiles_i * X.tiles_j * tiles_k : 0;
        X.tile0 = (int)tb.size();
        for (int t = 0; t < X.ntiles; ++t) tb.push_back(b);
    }
    B->ntiles = (int)tb.size();
    // eight contiguous queues of (nearly) equal length, block order kept
    for (int q = 0; q <= 8; ++q) B->qstart[q] = (int)((long)B->ntiles * q / 8);
    int rc;
    if (B->d_tile_block) { (void)hipFree(B->d_tile_block); B->d_tile_block = nullptr; }
    if (!tb.empty()) {
        if ((rc = check_hip(hipMalloc(&B->d_tile_block, tb.size() * sizeof(int32_t)), "hipMalloc"))) return rc;
        if ((rc = check_hip(hipMemcpy(B->d_tile_block, tb.data(), tb.size() * sizeof(int32_t),
                                      hipMemcpyHostToDevice), "hipMemcpy"))) return rc;
    }
    if ((rc = check_hip(hipMemcpy(B->db, B->hb.data(), B->hb.size() * sizeof(BBlock), hipMemcpyHostToDevice),
                        "hipMemcpy"))) return rc;
    B->tune.w = w;
    B->tune.ty = ty;
    B->tune.kc = kc;
    return LSM_OK;
}

template <bool FAST, int TY, int W>
int stage_tiles_t(LsmBatch *B, const QueueArgs &qa, hipStream_t s)
{
    const dim3 blk(TX, TY * W / TX, 1);
    if (B->tune.form == LSM_BATCH_QUEUE) {
        batch_queue_reset<<<1, 64, 0, s>>>(B->d_qcount);
        int per_cu = B->tune.wg_per_cu;
        if (per_cu <= 0) {
            int occ = 0;
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_stage_queue<FAST, TY, W>, TY * W, 0);
            per_cu = occ > 0 ? occ : 4;
        }
        unsigned grid = (unsigned)(B->cus * per_cu);
        if (grid > (unsigned)B->ntiles) grid = (unsigned)((B->ntiles + 7) / 8 * 8);
        batch_stage_queue<FAST, TY, W><<<grid, blk, 0, s>>>(qa);
    } else {
        const unsigned grid = (unsigned)((B->ntiles + 7) / 8 * 8);
        batch_stage_grid<FAST, TY, W><<<grid, blk, 0, s>>>(qa, B->ntiles);
    }
    return check_hip(hipGetLastError(), "batch stage");
}

template <bool FAST>
int stage_tiles(LsmBatch *B, const BFields &f, hipStream_t s)
{
    if (B->ntiles == 0) return LSM_OK;
    QueueArgs qa{};
    qa.blocks = B->db; qa.tile_block = B->d_tile_block; qa.qcount = B->d_qcount; qa.f = f;
    for (int q = 0; q <= 8; ++q) qa.qstart[q] = B->qstart[q];
    KTimer kt(LSM_KT_BATCH_STAGE, s);
    const int ty = B->tune.ty, w = B->tune.w;
    if (w == 32) return ty == 8 ? stage_tiles_t<FAST, 8, 32>(B, qa, s) : stage_tiles_t<FAST, 4, 32>(B, qa, s);
    return ty == 8 ? stage_tiles_t<FAST, 8, 64>(B, qa, s) : stage_tiles_t<FAST, 4, 64>(B, qa, s);
}

int shell_launch(LsmBatch *B, const BFields &f, int b0, int nb, hipStream_t s)
{
    const long base = B->hb[b0].shell0;
    const long end = (b0 + nb < B->nb) ? B->hb[b0 + nb].shell0 : B->nshell;
    if (end <= base) return LSM_OK;
    ShellBArgs sa{B->db, b0, nb, end - base, base, f};
    batch_shell<<<nblk(end - base), 256, 0, s>>>(sa);
    return check_hip(hipGetLastError(), "batch_shell");
}

int halo_launch(LsmBatch *B, double *field, int b0, int nb, hipStream_t s)
{
    if (!B->connected || B->g == 0) return LSM_OK;
    const int w0 = B->blk_wg[b0], w1 = B->blk_wg[b0 + nb];
    if (w1 <= w0) return LSM_OK;
    HaloArgs ha{};
    ha.blocks = B->db; ha.segs = B->d_segs; ha.wg_seg = B->d_wg_seg; ha.wg_base = w0;
    ha.field = field; ha.g = B->g;
    for (int ax = 0; ax < 3; ++ax) {
        ha.lat[ax] = B->lat[ax];
        for (int q = 0; q <= B->lat[ax]; ++q) ha.lstart[ax][q] = B->lstart[ax][q];
    }
    ha.block_of = B->d_block_of;
    batch_halo<<<(unsigned)(w1 - w0), 256, 0, s>>>(ha);
    return check_hip(hipGetLastError(), "batch_halo");
}

// halo segments (internal-face slabs) of every block and the workgroup table
int build_halo(LsmBatch *B)
{
    std::vector<HSeg> segs;
    std::vector<int32_t> wseg;
    B->blk_wg.assign(B->nb + 1, 0);
    for (int b = 0; b < B->nb; ++b) {
        B->blk_wg[b] = (int)wseg.size();
        if (!B->connected || B->g == 0) continue;
        const BBlock &X = B->hb[b];
        const int g = B->g;
        const int noi = X.r.ib - X.r.ia + 1, noj = X.r.jb - X.r.ja + 1, nok = X.r.kb - X.r.ka + 1;
        const bool internal[6] = {!X.r.fb, !X.r.ff, !X.r.fl, !X.r.fr, !X.r.fd, !X.r.fu};
        for (int sl = 0; sl < 6; ++sl) {
            if (!internal[sl]) continue;
            HSeg h{};
            h.b = b; h.slab = sl;
            if (sl < 2) {        // i slab: g x nj x nk
                h.n0 = g; h.n1 = X.nj; h.i0 = sl == 0 ? 0 : g + noi; h.j0 = 0; h.k0 = 0;
                h.count = g * X.nj * X.nk;
            } else if (sl < 4) { // j slab: noi x g x nk
                h.n0 = noi; h.n1 = g; h.i0 = g; h.j0 = sl == 2 ? 0 : g + noj; h.k0 = 0;
                h.count = noi * g * X.nk;
            } else {             // k slab: noi x noj x g
                h.n0 = noi; h.n1 = noj; h.i0 = g; h.j0 = g; h.k0 = sl == 4 ? 0 : g + nok;
                h.count = noi * noj * g;
            }
            h.wg0 = (int)wseg.size();
            const int nw = (h.count + 255) / 256;
            for (int q = 0; q < nw; ++q) wseg.push_back((int32_t)segs.size());
            segs.push_back(h);
        }
    }
    B->blk_wg[B->nb] = (int)wseg.size();
    B->halo_wgs = (int)wseg.size();
    if (segs.empty()) return LSM_OK;
    int rc;
    if ((rc = check_hip(hipMalloc(&B->d_segs, segs.size() * sizeof(HSeg)), "hipMalloc"))) return rc;
    if ((rc = check_hip(hipMalloc(&B->d_wg_seg, wseg.size() * sizeof(int32_t)), "hipMalloc"))) return rc;
    if ((rc = check_hip(hipMemcpy(B->d_segs, segs.data(), segs.size() * sizeof(HSeg), hipMemcpyHostToDevice),
                        "hipMemcpy"))) return rc;
    return check_hip(hipMemcpy(B->d_wg_seg, wseg.data(), wseg.size() * sizeof(int32_t), hipMemcpyHostToDevice),
                     "hipMemcpy");
}

template <class T>
int upload(const std::vector<T> &v, T **dst)
{
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    if (v.empty()) return LSM_OK;
    int rc = check_hip(hipMalloc(dst, v.size() * sizeof(T)), "hipMalloc");
    if (rc) return rc;
    return check_hip(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

// reinitialisation tiles (rt_w x 256/rt_w x 8 over each block's interior
// range), shell workgroups (256 physical-face nodes of one block each) and
// the bcond_lsm face segments of each axis
int build_reinit_tables(LsmBatch *B)
{
    drop_graph(B);
    const int W = B->rt_w, TY = 256 / W;
    std::vector<int32_t> rt, sh;
    std::vector<HSeg> bs[3];
    std::vector<int32_t> bw[3];
    for (int ax = 0; ax < 3; ++ax) B->bc_wg[ax].assign(B->nb + 1, 0);
    for (int b = 0; b < B->nb; ++b) {
        BBlock &X = B->hb[b];
        const Range &r = X.r;
        int kc = 8;
        if (const char *e = getenv("LSM_BATCH_REINIT_KCHUNK")) { int v = atoi(e); if (v > 0) kc = v; }
        X.rt_kc = kc;
        const int ni_int = r.ihi - r.ilo + 1, nj_int = r.jhi - r.jlo + 1, nk_int = r.khi - r.klo + 1;
        // first tile column on a 128-byte line at or before the first interior column
        X.rt_ib0 = max(0, r.ilo - (int)((X.off + r.ilo) % 16));
        X.rt_tiles_i = ni_int > 0 ? (r.ihi - X.rt_ib0 + 1 + W - 1) / W : 0;
        X.rt_tiles_j = nj_int > 0 ? (nj_int + TY - 1) / TY : 0;
        const int tk = nk_int > 0 ? (nk_int + kc - 1) / kc : 0;
        X.rt_ntiles = (r.ihi >= r.ilo) ? X.rt_tiles_i * X.rt_tiles_j * tk : 0;
        X.rt_tile0 = (int)rt.size();
        for (int t = 0; t < X.rt_ntiles; ++t) rt.push_back(b);
        X.sh_wg0 = (int)sh.size();
        const long nsh = X.nI + X.nJ + X.nK;
        for (long q = 0; q < (nsh + 255) / 256; ++q) sh.push_back(b);
        // bcond faces: axis ax, low (slab 0) / high (slab 1), owned extents of the other axes
        const int no[3] = {r.ib - r.ia + 1, r.jb - r.ja + 1, r.kb - r.ka + 1};
        const bool fl[3] = {r.fb, r.fl, r.fd}, fh[3] = {r.ff, r.fr, r.fu};
        for (int ax = 0; ax < 3; ++ax) {
            B->bc_wg[ax][b] = (int)bw[ax].size();
            const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
            for (int side = 0; side < 2; ++side) {
                if (!(side == 0 ? fl[ax] : fh[ax])) continue;
                HSeg h{};
                h.b = b; h.slab = side; h.n0 = no[o1]; h.n1 = no[o2];
                h.count = no[o1] * no[o2];
                h.wg0 = (int)bw[ax].size();
                for (int q = 0; q < (h.count + 255) / 256; ++q) bw[ax].push_back((int32_t)bs[ax].size());
                bs[ax].push_back(h);
            }
        }
    }
    for (int ax = 0; ax < 3; ++ax) B->bc_wg[ax][B->nb] = (int)bw[ax].size();
    B->rt_ntiles = (int)rt.size();
    B->sh_wgs = (int)sh.size();
    int rc;
    if ((rc = upload(rt, &B->d_rtile_block)) || (rc = upload(sh, &B->d_sh_block))) return rc;
    for (int ax = 0; ax < 3; ++ax)
        if ((rc = upload(bs[ax], &B->d_bseg[ax])) || (rc = upload(bw[ax], &B->d_bwg[ax]))) return rc;
    return check_hip(hipMemcpy(B->db, B->hb.data(), B->hb.size() * sizeof(BBlock), hipMemcpyHostToDevice),
                     "hipMemcpy");
}

int bcond_launch(LsmBatch *B, double *field, int b0, int nb, hipStream_t s)
{
    for (int ax = 0; ax < 3; ++ax) {
        const int w0 = B->bc_wg[ax][b0], w1 = B->bc_wg[ax][b0 + nb];
        if (w1 <= w0) continue;
        BcSegArgs ba{B->db, B->d_bseg[ax], B->d_bwg[ax], w0, ax, field};
        batch_bcond_phase<<<(unsigned)(w1 - w0), 256, 0, s>>>(ba);
        int rc = check_hip(hipGetLastError(), "batch_bcond_phase");
        if (rc) return rc;
    }
    return LSM_OK;
}

int ghost_copy_launch(LsmBatch *B, const double *src, double *dst, int b0, int nb, hipStream_t s)
{
    if (!B->connected || B->g == 0) return LSM_OK;
    const int w0 = B->blk_wg[b0], w1 = B->blk_wg[b0 + nb];
    if (w1 <= w0) return LSM_OK;
    HaloArgs ha{};
    ha.blocks = B->db; ha.segs = B->d_segs; ha.wg_seg = B->d_wg_seg; ha.wg_base = w0; ha.field = dst; ha.g = B->g;
    batch_ghost_copy<<<(unsigned)(w1 - w0), 256, 0, s>>>(ha, src);
    return check_hip(hipGetLastError(), "batch_ghost_copy");
}

template <int STAGE>
int reinit_stage_launch(LsmBatch *B, BRArgs a, int b0, int nb, hipStream_t s)
{
    const int t0 = B->hb[b0].rt_tile0;
    const int t1 = (b0 + nb < B->nb) ? B->hb[b0 + nb].rt_tile0 : B->rt_ntiles;
    const int w0 = B->hb[b0].sh_wg0;
    const int w1 = (b0 + nb < B->nb) ? B->hb[b0 + nb].sh_wg0 : B->sh_wgs;
    a.tile_base = t0; a.ntiles = t1 - t0; a.sh_base = w0;
    a.grid_tiles = (a.ntiles + 7) / 8 * 8;
    const unsigned grid = (unsigned)(a.grid_tiles + (w1 - w0));
    if (grid == 0) return LSM_OK;
    // the max / min reduction of the exit test is compiled in only where it is used
    constexpr bool CAN = STAGE == 3 || STAGE == 0;
    const bool red = CAN && a.reduce;
    if (red) {
        if (B->rt_w == 16) batch_reinit_stage<STAGE, 16, CAN><<<grid, dim3(16, 16, 1), 0, s>>>(a);
        else if (B->rt_w == 32) batch_reinit_stage<STAGE, 32, CAN><<<grid, dim3(32, 8, 1), 0, s>>>(a);
        else batch_reinit_stage<STAGE, 64, CAN><<<grid, dim3(64, 4, 1), 0, s>>>(a);
    } else {
        if (B->rt_w == 16) batch_reinit_stage<STAGE, 16, false><<<grid, dim3(16, 16, 1), 0, s>>>(a);
        else if (B->rt_w == 32) batch_reinit_stage<STAGE, 32, false><<<grid, dim3(32, 8, 1), 0, s>>>(a);
        else batch_reinit_stage<STAGE, 64, false><<<grid, dim3(64, 4, 1), 0, s>>>(a);
    }
    return check_hip(hipGetLastError(), "batch_reinit_stage");
}

size_t reinit_work_bytes(const LsmBatch *B)
{
    const size_t fb = ((size_t)B->total * sizeof(double) + 255) & ~(size_t)255;
    const size_t cb = ((size_t)B->total + 255) & ~(size_t)255;
    return 2 * fb + cb + (((size_t)B->nb * sizeof(RCtl) + 255) & ~(size_t)255) + 256;
}

const double kC[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};

}  // namespace

// per-block form: the single-block launches (lsm_advect.hip)
int launch_stage_block(const StageArgs &a, int ty, bool fast, hipStream_t s);
int launch_shell_block(const ShellArgs &sh, hipStream_t s);

int batch_rk3(LsmBatch *B, double *phi_np1, const double *phi_n, const double *U1, const double *U2,
              const double *U3, double dt, int mode, double *scratch, hipStream_t s)
{
    const double *src[3] = {phi_n, phi_np1, scratch};
    double *dst[3] = {phi_np1, scratch, phi_np1};
    const bool fast = mode != LSM_ADVECT_EXACT;
    int rc;
    for (int st = 0; st < 3; ++st) {
        BFields f{src[st], phi_n, U1, U2, U3, B->aj, dst[st], kC[st][0], kC[st][1], kC[st][2], dt};
        if (B->tune.form == LSM_BATCH_PERBLOCK) {
            for (int b = 0; b < B->nb; ++b) {
                const BBlock &X = B->hb[b];
                StageArgs a;
                a.ps = f.ps + X.off; a.pn = f.pn + X.off; a.out = f.out + X.off;
                a.U1 = U1 + X.off; a.U2 = U2 + X.off; a.U3 = U3 + X.off; a.aj = B->aj + X.off;
                a.mask = nullptr;
                a.ni = X.ni; a.nj = X.nj; a.nk = X.nk; a.sj = X.sj; a.sk = X.sk; a.r = X.r;
                a.c1 = f.c1; a.c2 = f.c2; a.c3 = f.c3; a.dt = dt; a.dc = X.dc; a.de = X.de; a.dz = X.dz;
                a.tiles_i = X.tiles_i; a.tiles_j = X.tiles_j; a.kc = X.kc; a.ntiles = X.ntiles; a.jb0 = X.jb0;
                a.ib0 = X.ib0;
                if (X.ntiles > 0 && (rc = launch_stage_block(a, B->tune.ty, fast, s))) return rc;
                ShellArgs sh{};
                sh.out_ro = f.out + X.off; sh.out = f.out + X.off; sh.ps = f.ps + X.off; sh.pn = f.pn + X.off;
                sh.aj = B->aj + X.off; sh.mask = nullptr;
                sh.ni = X.ni; sh.nj = X.nj; sh.nk = X.nk; sh.sj = X.sj; sh.sk = X.sk; sh.r = X.r;
                for (int q = 0; q < 6; ++q) sh.bt[q] = X.bt[q];
                sh.big = X.big; sh.c1 = f.c1; sh.c2 = f.c2; sh.c3 = f.c3; sh.dt = dt;
                sh.nI = X.nI; sh.nJ = X.nJ; sh.nK = X.nK; sh.cnt_i = X.cnt_i; sh.cnt_j = X.cnt_j;
                if ((rc = launch_shell_block(sh, s))) return rc;
            }
            for (int b = 0; b < B->nb; ++b)
                if ((rc = halo_launch(B, dst[st], b, 1, s))) return rc;
        } else {
            if (fast) { if ((rc = stage_tiles<true>(B, f, s))) return rc; }
            else if ((rc = stage_tiles<false>(B, f, s))) return rc;
            if ((rc = shell_launch(B, f, 0, B->nb, s))) return rc;
            if ((rc = halo_launch(B, dst[st], 0, B->nb, s))) return rc;
        }
    }
    return LSM_OK;
}

// reinitialisation_test2 on every block (ghost layers of phi0 kept), phi_out
// = result; blocks b0 .. b0+nb-1 per launch (all at once, or one by one in
// the per-block form)
int batch_reinit(LsmBatch *B, const LsmReinitParams &p, const double *phi0, double *out, int32_t *iters,
                 void *work, hipStream_t s)
{
    const size_t fb = ((size_t)B->total * sizeof(double) + 255) & ~(size_t)255;
    const size_t cb = ((size_t)B->total + 255) & ~(size_t)255;
    char *w = (char *)work;
    double *A = (double *)w, *Bf = (double *)(w + fb);
    uint8_t *conv = (uint8_t *)(w + 2 * fb);
    RCtl *ctl = (RCtl *)(w + 2 * fb + cb);
    const int ghosted = (B->connected && B->g > 0) ? 1 : 0;
    int rc;
    if ((rc = launch_copy(phi0, out, B->total, s))) return rc;
    if ((rc = ghost_copy_launch(B, phi0, A, 0, B->nb, s))) return rc;
    if ((rc = ghost_copy_launch(B, phi0, Bf, 0, B->nb, s))) return rc;
    batch_reinit_init<<<(B->nb + 63) / 64, 64, 0, s>>>(ctl, B->nb, ghosted, iters);
    if ((rc = check_hip(hipGetLastError(), "batch_reinit_init"))) return rc;
    if (p.rn_itermax <= 0) return LSM_OK;
    BRArgs a{};
    a.blocks = B->db; a.rtile_block = B->d_rtile_block; a.sh_block = B->d_sh_block;
    a.gt = B->gten; a.gs = B->total;
    a.phi0 = phi0; a.conv = conv; a.ctl = ctl;
    a.p = RStage{0, 0, 0, p.rn_timestep, p.epslsm, p.eps_reinit, p.order_bc, p.eno_bc};
    const bool per = B->tune.form == LSM_BATCH_PERBLOCK;
    const int nl = per ? B->nb : 1, span = per ? 1 : B->nb;
    // with ghost layers the minimum starts at 0, so the exit test can only
    // pass when eps_reinit > 1: otherwise every block runs rn_itermax
    // iterations and the max / min reductions are skipped
    a.reduce = (ghosted && !(p.eps_reinit > 1.0)) ? 0 : 1;
    a.din = phi0; a.dn = phi0; a.dout = nullptr;
    if (a.reduce)
        for (int q = 0; q < nl; ++q)
            if ((rc = reinit_stage_launch<0>(B, a, q * span, span, s))) return rc;
    const double *sin_[3] = {out, A, Bf};
    double *sout[3] = {A, Bf, out};
    for (int it = 0; it < p.rn_itermax; ++it) {
        batch_reinit_check<<<(B->nb + 63) / 64, 64, 0, s>>>(ctl, B->nb, ghosted, p.eps_reinit, iters);
        for (int st = 0; st < 3; ++st) {
            a.p = RStage{kC[st][0], kC[st][1], kC[st][2], p.rn_timestep, p.epslsm, p.eps_reinit, p.order_bc, p.eno_bc};
            a.din = sin_[st]; a.dn = out; a.dout = sout[st];
            KTimer kt(LSM_KT_BATCH_REINIT, s);
            for (int q = 0; q < nl; ++q) {
                if (st == 0) rc = reinit_stage_launch<1>(B, a, q * span, span, s);
                else if (st == 1) rc = reinit_stage_launch<2>(B, a, q * span, span, s);
                else rc = reinit_stage_launch<3>(B, a, q * span, span, s);
                if (rc) return rc;
            }
        }
    }
    return LSM_OK;
}

int batch_bcond_halo(LsmBatch *B, double *field, hipStream_t s)
{
    int rc;
    if (B->tune.form == LSM_BATCH_PERBLOCK) {
        for (int b = 0; b < B->nb; ++b)
            if ((rc = bcond_launch(B, field, b, 1, s))) return rc;
        for (int b = 0; b < B->nb; ++b)
            if ((rc = halo_launch(B, field, b, 1, s))) return rc;
        return LSM_OK;
    }
    if ((rc = bcond_launch(B, field, 0, B->nb, s))) return rc;
    return halo_launch(B, field, 0, B->nb, s);
}

// levelsetmethod.F90:221-326 for every block: ad_RKTVD3 (with bcond and halo
// fill after each stage), reinitialisation_test2, bcond_lsm, halo fill
int batch_step(LsmBatch *B, const LsmReinitParams *p, double *phi_np1, const double *phi_n, const double *U1,
               const double *U2, const double *U3, double dt, int mode, void *work, hipStream_t s)
{
    const size_t fb = ((size_t)B->total * sizeof(double) + 255) & ~(size_t)255;
    char *w = (char *)work;
    double *scr = (double *)w, *X = (double *)(w + fb);
    int rc;
    if (!p) return batch_rk3(B, phi_np1, phi_n, U1, U2, U3, dt, mode, scr, s);
    if ((rc = batch_rk3(B, X, phi_n, U1, U2, U3, dt, mode, scr, s))) return rc;
    if ((rc = batch_reinit(B, *p, X, phi_np1, nullptr, w + 2 * fb, s))) return rc;
    return batch_bcond_halo(B, phi_np1, s);
}

}  // namespace lsm

using namespace lsm;

extern "C" {

int lsm_batch_create(const LsmBatchDesc *d, LsmBatch **out)
{
    if (!d || !out) return fail(LSM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (d->nblocks < 1 || !d->blocks) return fail(LSM_ERR_ARG, "no blocks");
    if (d->ghost < 0 || d->ghost > 3) return fail(LSM_ERR_ARG, "ghost = %d (0..3)", d->ghost);
    if (!d->aj) return fail(LSM_ERR_ARG, "aj is NULL");
    const int g = d->ghost;
    const bool conn = d->connected != 0;
    if (conn && g < 2) return fail(LSM_ERR_ARG, "a connected batch needs ghost >= 2 (the WENO3 stencil)");
    LsmBatch *B = new LsmBatch();
    B->nb = d->nblocks; B->g = g; B->connected = conn ? 1 : 0; B->total = d->total;
    B->aj = d->aj;
    for (int c = 0; c < 3; ++c) { B->met[c] = d->csi[c]; B->met[3 + c] = d->eta[c]; B->met[6 + c] = d->zet[c]; }
    B->desc.assign(d->blocks, d->blocks + d->nblocks);
    auto bad = [&](int code, const char *msg, int b) {
        delete B;
        return fail(code, "block %d: %s", b, msg);
    };
    // lattice: positions, sizes per axis, global starts
    std::vector<int32_t> block_of;
    if (conn) {
        for (int ax = 0; ax < 3; ++ax) {
            B->lat[ax] = d->lattice[ax];
            if (B->lat[ax] < 1 || B->lat[ax] > kMaxLattice) { delete B; return fail(LSM_ERR_ARG, "lattice extent %d", B->lat[ax]); }
        }
        const long npos = (long)B->lat[0] * B->lat[1] * B->lat[2];
        if (npos != d->nblocks) { delete B; return fail(LSM_ERR_ARG, "a connected batch fills its lattice (%ld positions, %d blocks)", npos, d->nblocks); }
        block_of.assign(npos, -1);
        std::vector<int> sz[3];
        for (int ax = 0; ax < 3; ++ax) sz[ax].assign(B->lat[ax], -1);
        for (int b = 0; b < d->nblocks; ++b) {
            const LsmBatchBlock &X = d->blocks[b];
            const int n[3] = {X.ni, X.nj, X.nk};
            for (int ax = 0; ax < 3; ++ax) {
                const int L = X.lattice[ax];
                if (L < 0 || L >= B->lat[ax]) return bad(LSM_ERR_ARG, "lattice position outside the lattice", b);
                if (sz[ax][L] >= 0 && sz[ax][L] != n[ax]) return bad(LSM_ERR_ARG, "face sizes do not match its lattice slab", b);
                sz[ax][L] = n[ax];
            }
            const long p = X.lattice[0] + (long)B->lat[0] * (X.lattice[1] + (long)B->lat[1] * X.lattice[2]);
            if (block_of[p] >= 0) return bad(LSM_ERR_ARG, "two blocks at one lattice position", b);
            block_of[p] = b;
            // a face is physical exactly where the lattice ends
            int want = 0;
            if (X.lattice[0] == 0) want |= LSM_FACE_I1;
            if (X.lattice[0] == B->lat[0] - 1) want |= LSM_FACE_IM;
            if (X.lattice[1] == 0) want |= LSM_FACE_J1;
            if (X.lattice[1] == B->lat[1] - 1) want |= LSM_FACE_JM;
            if (X.lattice[2] == 0) want |= LSM_FACE_K1;
            if (X.lattice[2] == B->lat[2] - 1) want |= LSM_FACE_KM;
            if ((X.phys_faces & LSM_FACE_ALL) != want) return bad(LSM_ERR_ARG, "phys_faces do not match its lattice position", b);
        }
        for (int ax = 0; ax < 3; ++ax) {
            B->lstart[ax][0] = 0;
            for (int L = 0; L < B->lat[ax]; ++L) B->lstart[ax][L + 1] = B->lstart[ax][L] + sz[ax][L];
        }
    } else {
        if (g != 0) { delete B; return fail(LSM_ERR_ARG, "independent blocks carry no ghost layers"); }
        for (int b = 0; b < d->nblocks; ++b)
            if ((d->blocks[b].phys_faces & LSM_FACE_ALL) != LSM_FACE_ALL)
                return bad(LSM_ERR_ARG, "an independent block has six physical faces", b);
        B->lat[0] = B->lat[1] = B->lat[2] = 1;
    }
    // block descriptors
    long shell = 0;
    B->hb.resize(d->nblocks);
    for (int b = 0; b < d->nblocks; ++b) {
        const LsmBatchBlock &X = d->blocks[b];
        if (X.ni < 4 || X.nj < 4 || X.nk < 4) return bad(LSM_ERR_ARG, "every extent must be >= 4", b);
        for (int q = 0; q < 6; ++q)
            if (X.btype[q] < 0 || X.btype[q] > 2) return bad(LSM_ERR_ARG, "btype outside 0..2", b);
        BBlock &Y = B->hb[b];
        memset(&Y, 0, sizeof Y);
        Y.off = X.offset;
        Y.ni = X.ni + 2 * g; Y.nj = X.nj + 2 * g; Y.nk = X.nk + 2 * g;
        if (X.pitch != 0 && X.pitch < Y.ni) return bad(LSM_ERR_ARG, "row pitch below ni + 2 ghost", b);
        Y.sj = X.pitch ? X.pitch : Y.ni;
        Y.sk = Y.sj * Y.nj;
        if (Y.off < 0 || Y.off + Y.sk * Y.nk > d->total) return bad(LSM_ERR_ARG, "array outside the batch arrays", b);
        Y.r = block_range(X, g);
        for (int ax = 0; ax < 3; ++ax) Y.gs[ax] = conn ? B->lstart[ax][X.lattice[ax]] : 0;
        for (int q = 0; q < 6; ++q) Y.bt[q] = X.btype[q];
        Y.big = X.big_phi; Y.dc = X.dc; Y.de = X.de; Y.dz = X.dz;
        ShellArgs sh{};
        sh.r = Y.r;
        shell_counts(Y.r, sh);
        Y.nI = sh.nI; Y.nJ = sh.nJ; Y.nK = sh.nK; Y.cnt_i = sh.cnt_i; Y.cnt_j = sh.cnt_j;
        Y.shell0 = shell;
        shell += sh.nI + sh.nJ + sh.nK;
    }
    B->nshell = shell;
    int rc;
    if ((rc = check_hip(hipGetDevice(&B->dev), "hipGetDevice"))) { delete B; return rc; }
    if ((rc = check_hip(hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, B->dev),
                        "hipDeviceGetAttribute"))) { delete B; return rc; }
    if ((rc = check_hip(hipMalloc(&B->db, B->hb.size() * sizeof(BBlock)), "hipMalloc")) ||
        (rc = check_hip(hipMalloc(&B->d_qcount, 64 * sizeof(int)), "hipMalloc"))) {
        lsm_batch_destroy(B);
        return rc;
    }
    if (conn) {
        if ((rc = check_hip(hipMalloc(&B->d_block_of, block_of.size() * sizeof(int32_t)), "hipMalloc")) ||
            (rc = check_hip(hipMemcpy(B->d_block_of, block_of.data(), block_of.size() * sizeof(int32_t),
                                      hipMemcpyHostToDevice), "hipMemcpy"))) {
            lsm_batch_destroy(B);
            return rc;
        }
    }
    if ((rc = build_tiles(B, 64, 4, 8))) { lsm_batch_destroy(B); return rc; }
    if ((rc = build_halo(B))) { lsm_batch_destroy(B); return rc; }
    if ((rc = build_reinit_tables(B))) { lsm_batch_destroy(B); return rc; }
    // the reinitialisation reads the metric tensor (computed here once: the
    // batch's grid is fixed for its lifetime)
    bool have_met = true;
    for (int c = 0; c < 9; ++c) have_met = have_met && B->met[c];
    if (have_met) {
        if ((rc = check_hip(hipMalloc(&B->gten, 6 * (size_t)B->total * sizeof(double)), "hipMalloc")) ||
            (rc = launch_metric_tensor(B->met, B->total, B->gten, nullptr)) ||
            (rc = check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize"))) {
            lsm_batch_destroy(B);
            return rc;
        }
    }
    *out = B;
    return LSM_OK;
}

int lsm_batch_destroy(LsmBatch *B)
{
    if (!B) return LSM_OK;
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->db) (void)hipFree(B->db);
    if (B->d_tile_block) (void)hipFree(B->d_tile_block);
    if (B->d_block_of) (void)hipFree(B->d_block_of);
    if (B->d_segs) (void)hipFree(B->d_segs);
    if (B->d_rtile_block) (void)hipFree(B->d_rtile_block);
    if (B->d_sh_block) (void)hipFree(B->d_sh_block);
    if (B->gten) (void)hipFree(B->gten);
    for (int ax = 0; ax < 3; ++ax) {
        if (B->d_bseg[ax]) (void)hipFree(B->d_bseg[ax]);
        if (B->d_bwg[ax]) (void)hipFree(B->d_bwg[ax]);
    }
    if (B->d_wg_seg) (void)hipFree(B->d_wg_seg);
    if (B->d_qcount) (void)hipFree(B->d_qcount);
    delete B;
    return LSM_OK;
}

int lsm_batch_set_config(LsmBatch *B, int form, int tile_cols, int tile_rows, int kchunk, int wg_per_cu)
{
    if (!B) return fail(LSM_ERR_ARG, "NULL batch");
    if (form != LSM_BATCH_QUEUE && form != LSM_BATCH_GRID && form != LSM_BATCH_PERBLOCK)
        return fail(LSM_ERR_ARG, "form %d", form);
    if (tile_cols != 64 && tile_cols != 32) return fail(LSM_ERR_ARG, "tile_cols %d (64 or 32)", tile_cols);
    if (form == LSM_BATCH_PERBLOCK && tile_cols != 64)
        return fail(LSM_ERR_ARG, "the per-block form runs the single-block kernels (tile_cols 64)");
    if (tile_rows != 4 && tile_rows != 8) return fail(LSM_ERR_ARG, "tile_rows %d (4 or 8)", tile_rows);
    if (kchunk < 1 || kchunk > 64) return fail(LSM_ERR_ARG, "kchunk %d (1..64)", kchunk);
    if (wg_per_cu < 0 || wg_per_cu > 32) return fail(LSM_ERR_ARG, "wg_per_cu %d (0..32)", wg_per_cu);
    int rc;
    if (tile_cols != B->tune.w || tile_rows != B->tune.ty || kchunk != B->tune.kc) {
        // the device tables may still be read by queued work
        if ((rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) return rc;
        if ((rc = build_tiles(B, tile_cols, tile_rows, kchunk))) return rc;
    } else if (B->gexec && (form != B->tune.form || wg_per_cu != B->tune.wg_per_cu)) {
        // a replay may be in flight
        if ((rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) return rc;
        drop_graph(B);
    }
    B->tune.form = form;
    B->tune.wg_per_cu = wg_per_cu;
    return LSM_OK;
}

int lsm_batch_set_reinit_tile(LsmBatch *B, int tile_cols)
{
    if (!B) return fail(LSM_ERR_ARG, "NULL batch");
    if (tile_cols != 0 && tile_cols != 64 && tile_cols != 32 && tile_cols != 16)
        return fail(LSM_ERR_ARG, "tile_cols %d (0, 64, 32 or 16)", tile_cols);
    int rc;
    // the device tables may still be read by queued work
    if ((rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) return rc;
    B->rt_w = tile_cols ? tile_cols : 64;
    return build_reinit_tables(B);
}

int lsm_batch_get_reinit_tile(const LsmBatch *B)
{
    return B ? B->rt_w : 0;
}

int lsm_batch_get_config(const LsmBatch *B, int32_t cfg[5])
{
    if (!B || !cfg) return fail(LSM_ERR_ARG, "NULL argument");
    cfg[0] = B->tune.form; cfg[1] = B->tune.w; cfg[2] = B->tune.ty; cfg[3] = B->tune.kc;
    cfg[4] = B->tune.wg_per_cu;
    return LSM_OK;
}

size_t lsm_batch_workspace_bytes(const LsmBatch *B, int op)
{
    if (!B) return 0;
    const size_t fb = ((size_t)B->total * sizeof(double) + 255) & ~(size_t)255;
    if (op == LSM_OP_RK3) return fb;
    if (op == LSM_OP_REINIT_ENO2) return reinit_work_bytes(B);
    if (op == LSM_OP_LEVELSET_STEP) return 2 * fb + reinit_work_bytes(B);
    return 0;
}

int lsm_batch_halo_fill(LsmBatch *B, double *field, void *stream)
{
    if (!B || !field) return fail(LSM_ERR_ARG, "NULL argument");
    return halo_launch(B, field, 0, B->nb, (hipStream_t)stream);
}

int lsm_batch_contra_velocity(LsmBatch *B, const double *u, const double *v, const double *w, double *ucn1,
                              double *ucn2, double *ucn3, void *stream)
{
    if (!B || !u || !v || !w || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL argument");
    Blk b{};
    b.csi0 = B->met[0]; b.csi1 = B->met[1]; b.csi2 = B->met[2];
    b.eta0 = B->met[3]; b.eta1 = B->met[4]; b.eta2 = B->met[5];
    b.zet0 = B->met[6]; b.zet1 = B->met[7]; b.zet2 = B->met[8];
    b.aj = B->aj; b.n = B->total;
    for (int q = 0; q < 9; ++q)
        if (!(&b.csi0)[q]) return fail(LSM_ERR_ARG, "the batch was created without metrics");
    return launch_contra(b, u, v, w, 1, ucn1, ucn2, ucn3, (hipStream_t)stream);
}

int lsm_batch_rk3_step(LsmBatch *B, double *phi_np1, const double *phi_n, const double *ucn1, const double *ucn2,
                       const double *ucn3, double delti_lsm, int mode, void *work, size_t work_bytes, void *stream)
{
    if (!B || !phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL argument");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (mode == LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 (LSM_ADVECT_UPWIND1): the batched form covers the WENO3 right-hand side only");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    const size_t need = lsm_batch_workspace_bytes(B, LSM_OP_RK3);
    void *scr = work;
    if (!scr) {
        int rc = default_work(need, (hipStream_t)stream, &scr);
        if (rc) return rc;
    } else if (work_bytes < need) {
        return fail(LSM_ERR_WORKSPACE, "workspace of %zu bytes, %zu needed", work_bytes, need);
    }
    return batch_rk3(B, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, (double *)scr, (hipStream_t)stream);
}

// ---- autotuner ----------------------------------------------------------
// Candidates: the queue form with tile rows 4 / 8, 4 / 8 / 16 planes per tile
// and as many persistent workgroups per CU as fit or 4; the grid form with the
// same tiles.  Each is timed over `reps` steps (after one untimed step) on the
// caller's fields; the fastest is kept and remembered for the batch's shape
// signature (block extents, ghost layers, arithmetic), so a second batch of the
// same shapes takes it without timing.
static std::mutex g_tune_mu;
static std::map<std::vector<long>, BatchTune> g_tune_cache;

static std::vector<long> tune_key(const LsmBatch *B, int mode)
{
    std::vector<long> k{B->nb, B->g, B->connected, mode};
    for (const auto &d : B->desc) { k.push_back(d.ni); k.push_back(d.nj); k.push_back(d.nk); }
    return k;
}

int lsm_batch_autotune(LsmBatch *B, double *phi_np1, const double *phi_n, const double *ucn1, const double *ucn2,
                       const double *ucn3, double delti_lsm, int mode, void *work, size_t work_bytes, int reps,
                       int force, double *best_ms, void *stream)
{
    if (!B || !phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3 || !work) return fail(LSM_ERR_ARG, "NULL argument");
    if (reps < 1) return fail(LSM_ERR_ARG, "reps must be >= 1");
    const std::vector<long> key = tune_key(B, mode);
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            const BatchTune t = it->second;
            if (best_ms) *best_ms = -1.0;
            return lsm_batch_set_config(B, t.form, t.w, t.ty, t.kc, t.wg_per_cu);
        }
    }
    hipStream_t s = (hipStream_t)stream;
    std::vector<BatchTune> cand;
    for (int w : {64, 32})
        for (int ty : {4, 8})
            for (int kc : {4, 8, 16}) {
                cand.push_back(BatchTune{LSM_BATCH_QUEUE, w, ty, kc, 0});
                cand.push_back(BatchTune{LSM_BATCH_QUEUE, w, ty, kc, 4});
                cand.push_back(BatchTune{LSM_BATCH_GRID, w, ty, kc, 0});
            }
    hipEvent_t e0, e1;
    int rc;
    if ((rc = check_hip(hipEventCreate(&e0), "hipEventCreate"))) return rc;
    if ((rc = check_hip(hipEventCreate(&e1), "hipEventCreate"))) { (void)hipEventDestroy(e0); return rc; }
    double best = 1e300;
    BatchTune win = B->tune;
    for (const BatchTune &c : cand) {
        if ((rc = lsm_batch_set_config(B, c.form, c.w, c.ty, c.kc, c.wg_per_cu))) break;
        if ((rc = lsm_batch_rk3_step(B, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, work, work_bytes, stream)))
            break;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && !rc; ++r)
            rc = lsm_batch_rk3_step(B, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, work, work_bytes, stream);
        if (rc) break;
        (void)hipEventRecord(e1, s);
        if ((rc = check_hip(hipEventSynchronize(e1), "hipEventSynchronize"))) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / reps;
        if (per < best) { best = per; win = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = win;
    }
    if (best_ms) *best_ms = best;
    return lsm_batch_set_config(B, win.form, win.w, win.ty, win.kc, win.wg_per_cu);
}

// ---- graph of a full step -------------------------------------------------
static int check_step_args(LsmBatch *B, const LsmReinitParams *p, double *phi_np1, const double *phi_n,
                           const double *ucn1, const double *ucn2, const double *ucn3, int mode, void *work,
                           size_t work_bytes)
{
    if (!B || !phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL argument");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (mode == LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 (LSM_ADVECT_UPWIND1): the batched form covers the WENO3 right-hand side only");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    if (p) {
        if (p->hybrid)
            return fail(LSM_ERR_UNSUPPORTED, "hybrid (geometric) reinitialisation: the batched form covers the Sussman path only");
        if (p->order_bc < 1 || p->order_bc > 3) return fail(LSM_ERR_ARG, "order_bc = %d (1, 2 or 3)", p->order_bc);
        if (p->rn_itermax < 0) return fail(LSM_ERR_ARG, "rn_itermax < 0");
        if (!(p->epslsm > 0.0)) return fail(LSM_ERR_ARG, "epslsm must be > 0");
        for (int c = 0; c < 9; ++c)
            if (!B->met[c]) return fail(LSM_ERR_ARG, "the batch was created without metrics");
    }
    if (!work) return fail(LSM_ERR_ARG, "NULL workspace");
    const size_t need = lsm_batch_workspace_bytes(B, p ? LSM_OP_LEVELSET_STEP : LSM_OP_RK3);
    if (work_bytes < need) return fail(LSM_ERR_WORKSPACE, "workspace of %zu bytes, %zu needed", work_bytes, need);
    return LSM_OK;
}

int lsm_batch_reinit_sussman_eno2(LsmBatch *B, const LsmReinitParams *p, const double *phi0, double *phi_out,
                                  int32_t *iters, void *work, size_t work_bytes, void *stream)
{
    if (!B || !p || !phi0 || !phi_out || !work) return fail(LSM_ERR_ARG, "NULL argument");
    if (phi0 == phi_out) return fail(LSM_ERR_ARG, "phi0 and phi_out must not alias");
    if (p->order_bc < 1 || p->order_bc > 3) return fail(LSM_ERR_ARG, "order_bc = %d (1, 2 or 3)", p->order_bc);
    if (p->rn_itermax < 0) return fail(LSM_ERR_ARG, "rn_itermax < 0");
    if (!(p->epslsm > 0.0)) return fail(LSM_ERR_ARG, "epslsm must be > 0");
    for (int c = 0; c < 9; ++c)
        if (!B->met[c]) return fail(LSM_ERR_ARG, "the batch was created without metrics");
    if (work_bytes < reinit_work_bytes(B)) return fail(LSM_ERR_WORKSPACE, "workspace too small");
    return batch_reinit(B, *p, phi0, phi_out, iters, work, (hipStream_t)stream);
}

int lsm_batch_bcond(LsmBatch *B, double *field, void *stream)
{
    if (!B || !field) return fail(LSM_ERR_ARG, "NULL argument");
    return bcond_launch(B, field, 0, B->nb, (hipStream_t)stream);
}

int lsm_batch_levelset_step(LsmBatch *B, const LsmReinitParams *p, double *phi_np1, const double *phi_n,
                            const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm, int mode,
                            void *work, size_t work_bytes, void *stream)
{
    int rc = check_step_args(B, p, phi_np1, phi_n, ucn1, ucn2, ucn3, mode, work, work_bytes);
    if (rc) return rc;
    return batch_step(B, p, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, work, (hipStream_t)stream);
}

int lsm_batch_graph_capture(LsmBatch *B, const LsmReinitParams *p, double *phi_np1, const double *phi_n,
                            const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm, int mode,
                            void *work, size_t work_bytes)
{
    int rc0 = check_step_args(B, p, phi_np1, phi_n, ucn1, ucn2, ucn3, mode, work, work_bytes);
    if (rc0) return rc0;
    if (timing_on()) return fail(LSM_ERR_UNSUPPORTED, "switch launch timing off before a graph capture");
    int rc;
    hipStream_t cs;
    // a replay of the previous graph may be in flight
    if (B->gexec && (rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) return rc;
    drop_graph(B);
    if ((rc = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    rc = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    int rc2 = LSM_OK;
    if (!rc) {
        rc2 = batch_step(B, p, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, work, cs);
        hipGraph_t g = nullptr;
        rc = check_hip(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
        if (!rc && !rc2) {
            B->graph = g;
            rc = check_hip(hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
        } else if (g) {
            (void)hipGraphDestroy(g);
        }
    }
    (void)hipStreamDestroy(cs);
    return rc2 ? rc2 : rc;
}

int lsm_batch_graph_launch(LsmBatch *B, void *stream)
{
    if (!B) return fail(LSM_ERR_ARG, "NULL batch");
    if (!B->gexec) return fail(LSM_ERR_ARG, "no graph captured");
    return check_hip(hipGraphLaunch(B->gexec, (hipStream_t)stream), "hipGraphLaunch");
}

}  // extern "C"
