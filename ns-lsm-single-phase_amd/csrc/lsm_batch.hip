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
    int tiles_i, tiles_j, ntiles, tile0, kc, jb0;
    long shell0;              // first shell node of the block in the batch enumeration
    long nI, nJ, nK;
    int cnt_i, cnt_j;
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
    a.tiles_i = B.tiles_i; a.tiles_j = B.tiles_j; a.kc = B.kc; a.ntiles = B.ntiles; a.jb0 = B.jb0;
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

// (re)builds the tile tables for tiles of w columns, ty rows and kc planes
int build_tiles(LsmBatch *B, int w, int ty, int kc)
{
    std::vector<int32_t> tb;
    for (int b = 0; b < B->nb; ++b) {
        BBlock &X = B->hb[b];
        const Range &r = X.r;
        X.tiles_i = (r.ihi + 1 + w - 1) / w;
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
        Y.sj = Y.ni; Y.sk = (long)Y.ni * Y.nj;
        if (Y.off < 0 || Y.off + (long)Y.ni * Y.nj * Y.nk > d->total) return bad(LSM_ERR_ARG, "array outside the batch arrays", b);
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
    }
    B->tune.form = form;
    B->tune.wg_per_cu = wg_per_cu;
    return LSM_OK;
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
    if (op == LSM_OP_RK3) return ((size_t)B->total * sizeof(double) + 255) & ~(size_t)255;
    return 0;
}

int lsm_batch_halo_fill(LsmBatch *B, double *field, void *stream)
{
    if (!B || !field) return fail(LSM_ERR_ARG, "NULL argument");
    return halo_launch(B, field, 0, B->nb, (hipStream_t)stream);
}

int lsm_batch_contra_velocity(LsmBatch *B, const double *csi[3], const double *eta[3], const double *zet[3],
                              const double *u, const double *v, const double *w, double *ucn1, double *ucn2,
                              double *ucn3, void *stream)
{
    if (!B || !csi || !eta || !zet || !u || !v || !w || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL argument");
    Blk b{};
    b.csi0 = csi[0]; b.csi1 = csi[1]; b.csi2 = csi[2];
    b.eta0 = eta[0]; b.eta1 = eta[1]; b.eta2 = eta[2];
    b.zet0 = zet[0]; b.zet1 = zet[1]; b.zet2 = zet[2];
    b.aj = B->aj; b.n = B->total;
    for (int q = 0; q < 9; ++q)
        if (!(&b.csi0)[q]) return fail(LSM_ERR_ARG, "NULL metric");
    return launch_contra(b, u, v, w, 1, ucn1, ucn2, ucn3, (hipStream_t)stream);
}

int lsm_batch_rk3_step(LsmBatch *B, double *phi_np1, const double *phi_n, const double *ucn1, const double *ucn2,
                       const double *ucn3, double delti_lsm, int mode, void *work, size_t work_bytes, void *stream)
{
    if (!B || !phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL argument");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    const size_t need = lsm_batch_workspace_bytes(B, LSM_OP_RK3);
    void *scr = work;
    if (!scr) {
        int rc = default_work(need, &scr);
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
int lsm_batch_graph_capture(LsmBatch *B, double *phi_np1, const double *phi_n, const double *ucn1,
                            const double *ucn2, const double *ucn3, double delti_lsm, int mode, void *work,
                            size_t work_bytes)
{
    if (!B || !phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3 || !work) return fail(LSM_ERR_ARG, "NULL argument");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    if (work_bytes < lsm_batch_workspace_bytes(B, LSM_OP_RK3)) return fail(LSM_ERR_WORKSPACE, "workspace too small");
    if (timing_on()) return fail(LSM_ERR_UNSUPPORTED, "switch launch timing off before a graph capture");
    int rc;
    hipStream_t cs;
    if ((rc = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    if (B->gexec) { (void)hipGraphExecDestroy(B->gexec); B->gexec = nullptr; }
    if (B->graph) { (void)hipGraphDestroy(B->graph); B->graph = nullptr; }
    rc = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    int rc2 = LSM_OK;
    if (!rc) {
        rc2 = batch_rk3(B, phi_np1, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, (double *)work, cs);
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
