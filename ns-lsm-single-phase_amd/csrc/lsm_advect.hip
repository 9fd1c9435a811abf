// WENO3 + SSP-RK3 level-set advection (ad_RKTVD3 / calc_RH_AD_WENO3), the
// contravariant velocity (rhs_contra_j) and the level-set boundary conditions
// (bcond_lsm) for one block resident in HBM.
//
// Hot kernel: advect_stage.  One workgroup owns a 64 (i) x 4 (j) column tile
// and marches up a chunk of k planes.  Each thread keeps phi(k-2..k+2) of its
// own column in registers; the current k plane of the tile, with the two-node
// i/j halo the WENO3 stencil needs, is staged in LDS (double buffered, one
// barrier per plane).  Each stage reads phi once, the stage-0 field phi^n, the
// three U/J components and aj at the centre, and writes the new phi: the
// whole stage update (RHS + RK combination) is fused.  The boundary shell is
// written by a second, tiny kernel that applies the RK combination and the
// face extrapolations of bcond_lsm in the reference's face order.
#include <stdarg.h>
#include <stdio.h>

#include "lsm_internal.hpp"
#include "lsm_advect_tile.hpp"

namespace lsm {

namespace {

template <bool FAST, bool MASK, int TY, bool NT = true, bool FOLD = false>
__global__ __launch_bounds__(TX * TY) void advect_stage(StageArgs a)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.ntiles) return;
    advect_tile<FAST, MASK, TY, NT, FOLD>(a, tile);
}

template <bool MASK>
__global__ __launch_bounds__(256) void stage_shell(ShellArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!shell_node(a, t, i, j, k)) return;
    const double v = sh_f3<MASK>(a, i, j, k);
    a.out[(long)i + a.sj * j + a.sk * k] = v;
}

// ---------------------------------------------------------------------------
// standalone bcond_lsm: three in-place phases (i faces, j faces, k faces)
// ---------------------------------------------------------------------------
struct BcArgs {
    double *phi;
    int ni, nj, nk;
    long sj, sk;
    Range r;
    int bt[6];
    double big;
};

__global__ __launch_bounds__(256) void bcond_phase(BcArgs a, int axis)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n1, n2;
    long s, s1, s2;
    bool f_lo, f_hi;
    int lo, hi, bt_lo, bt_hi;
    if (axis == 0) { n1 = a.nj; n2 = a.nk; s = 1; s1 = a.sj; s2 = a.sk; f_lo = a.r.fb; f_hi = a.r.ff;
                     lo = a.r.ia; hi = a.r.ib; bt_lo = a.bt[0]; bt_hi = a.bt[1]; }
    else if (axis == 1) { n1 = a.ni; n2 = a.nk; s = a.sj; s1 = 1; s2 = a.sk; f_lo = a.r.fl; f_hi = a.r.fr;
                          lo = a.r.ja; hi = a.r.jb; bt_lo = a.bt[2]; bt_hi = a.bt[3]; }
    else { n1 = a.ni; n2 = a.nj; s = a.sk; s1 = 1; s2 = a.sj; f_lo = a.r.fd; f_hi = a.r.fu;
           lo = a.r.ka; hi = a.r.kb; bt_lo = a.bt[4]; bt_hi = a.bt[5]; }
    const long per = (long)n1 * n2;
    if (t >= 2 * per) return;
    const int f = (int)(t / per);
    const long rem = t % per;
    const long base = (rem % n1) * s1 + (rem / n1) * s2;
    if (f == 0 && f_lo) {
        const long l = base + lo * s;
        if (bt_lo != 2) a.phi[l] = bc_face(bt_lo, a.phi[l + s], a.phi[l + 2 * s], a.phi[l], a.big, false);
    } else if (f == 1 && f_hi) {
        const long l = base + hi * s;
        a.phi[l] = bc_face(bt_hi, a.phi[l - s], a.phi[l - 2 * s], a.phi[l], a.big, axis == 0);
    }
}

// ---------------------------------------------------------------------------
// standalone WENO3 right-hand side (calc_RH_AD_WENO3) on interior nodes
// ---------------------------------------------------------------------------
struct RhsArgs {
    const double *phi, *U1, *U2, *U3;
    const int32_t *mask;
    double *rhs;
    int ni, nj, nk;
    long sj, sk;
    Range r;
    double dc, de, dz;
};

template <bool FAST>
__global__ __launch_bounds__(256) void weno3_rhs_kernel(RhsArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + a.r.ilo;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + a.r.jlo;
    const int k = blockIdx.z + a.r.klo;
    if (i > a.r.ihi || j > a.r.jhi || k > a.r.khi) return;
    const long l = (long)i + a.sj * j + a.sk * k;
    if (a.mask && a.mask[l] == 0) return;
    const double *p = a.phi;
    const long sj = a.sj, sk = a.sk;
    auto at = [&](long q) -> double { return p[q]; };
    // out-of-range stencil entries are only read when the stencil is not
    // forced; forced stencils replace them by a dummy, so guard the reads
    const bool fmb = a.r.fb && i == a.r.ilo, fmf = a.r.ff && i == a.r.ihi;
    const bool fml = a.r.fl && j == a.r.jlo, fmr = a.r.fr && j == a.r.jhi;
    const bool fmd = a.r.fd && k == a.r.klo, fmu = a.r.fu && k == a.r.khi;
    const double xm2 = fmb ? 0.0 : at(l - 2), xp2 = fmf ? 0.0 : at(l + 2);
    const double ym2 = fml ? 0.0 : at(l - 2 * sj), yp2 = fmr ? 0.0 : at(l + 2 * sj);
    const double zm2 = fmd ? 0.0 : at(l - 2 * sk), zp2 = fmu ? 0.0 : at(l + 2 * sk);
    const double c = at(l);
    const double Lc = weno3_dir<FAST>(xm2, at(l - 1), c, at(l + 1), xp2, a.U1[l], fmb, fmf, a.dc);
    const double Le = weno3_dir<FAST>(ym2, at(l - sj), c, at(l + sj), yp2, a.U2[l], fml, fmr, a.de);
    const double Lz = weno3_dir<FAST>(zm2, at(l - sk), c, at(l + sk), zp2, a.U3[l], fmd, fmu, a.dz);
    a.rhs[l] = (Lc + Le) + Lz;
}

// ---------------------------------------------------------------------------
// rhs_contra_j: ucn_j = (1/aj) * (grad xi^j . u)
// ---------------------------------------------------------------------------
struct ContraArgs {
    const double *u, *v, *w;
    long stride;
    const double *c0, *c1, *c2, *e0, *e1, *e2, *z0, *z1, *z2, *aj;
    double *U1, *U2, *U3;
    long n;
};

__global__ __launch_bounds__(256) void contra_kernel(ContraArgs a)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.n) return;
    const double u = a.u[l * a.stride], v = a.v[l * a.stride], w = a.w[l * a.stride];
    const double tmp = 1.0 / a.aj[l];
    a.U1[l] = tmp * ((a.c0[l] * u + a.c1[l] * v) + a.c2[l] * w);
    a.U2[l] = tmp * ((a.e0[l] * u + a.e1[l] * v) + a.e2[l] * w);
    a.U3[l] = tmp * ((a.z0[l] * u + a.z1[l] * v) + a.z2[l] * w);
}

// frozen-velocity operator of the fast path: V_d = (aj * dt * d) * ucn_d;
// counts nodes where aj is not finite and positive (for aj <= 0 sign(V) would
// not be the upwind sign)
__global__ __launch_bounds__(256) void fold_kernel(long n, const double *__restrict__ aj,
                                                   const double *__restrict__ U1,
                                                   const double *__restrict__ U2,
                                                   const double *__restrict__ U3, double dt, double dc,
                                                   double de, double dz, double *__restrict__ V1,
                                                   double *__restrict__ V2, double *__restrict__ V3,
                                                   int *bad)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const double a = aj[l];
    if (!(a > 0.0) || __builtin_isinf(a)) atomicAdd(bad, 1);
    const double s = a * dt;
    V1[l] = (s * dc) * U1[l];
    V2[l] = (s * de) * U2[l];
    V3[l] = (s * dz) * U3[l];
}

__global__ __launch_bounds__(256) void copy_kernel(const double *__restrict__ s, double *__restrict__ d, long n)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) d[l] = s[l];
}

inline unsigned blocks_for(long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// one block's stage tiles and shell with the single-block kernels (the
// per-block form of the batched step)
int launch_stage_block(const StageArgs &a, int ty, bool fast, hipStream_t s)
{
    const unsigned grid = (unsigned)((a.ntiles + 7) / 8 * 8);
    const dim3 blk(TX, ty, 1);
    if (ty == 8) {
        if (fast) advect_stage<true, false, 8><<<grid, blk, 0, s>>>(a);
        else advect_stage<false, false, 8><<<grid, blk, 0, s>>>(a);
    } else {
        if (fast) advect_stage<true, false, 4><<<grid, blk, 0, s>>>(a);
        else advect_stage<false, false, 4><<<grid, blk, 0, s>>>(a);
    }
    return check_hip(hipGetLastError(), "advect_stage");
}

int launch_shell_block(const ShellArgs &sh, hipStream_t s)
{
    const long n = sh.nI + sh.nJ + sh.nK;
    if (n <= 0) return LSM_OK;
    stage_shell<false><<<blocks_for(n, 256), 256, 0, s>>>(sh);
    return check_hip(hipGetLastError(), "stage_shell");
}

int launch_copy(const double *src, double *dst, long n, hipStream_t s)
{
    if (n <= 0) return LSM_OK;
    copy_kernel<<<blocks_for(n, 256), 256, 0, s>>>(src, dst, n);
    return check_hip(hipGetLastError(), "copy_kernel");
}

int launch_contra(const Blk &b, const double *u, const double *v, const double *w, long stride,
                  double *U1, double *U2, double *U3, hipStream_t s)
{
    ContraArgs a{u, v, w, stride, b.csi0, b.csi1, b.csi2, b.eta0, b.eta1, b.eta2,
                 b.zet0, b.zet1, b.zet2, b.aj, U1, U2, U3, b.n};
    contra_kernel<<<blocks_for(b.n, 256), 256, 0, s>>>(a);
    return check_hip(hipGetLastError(), "contra_kernel");
}

int launch_weno3_rhs(const Blk &b, const double *phi, const double *U1, const double *U2,
                     const double *U3, const int32_t *mask, double *rhs, int mode, hipStream_t s)
{
    const Range r = make_range(b);
    RhsArgs a{phi, U1, U2, U3, mask, rhs, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, r, b.dc, b.de, b.dz};
    const int ni = r.ihi - r.ilo + 1, nj = r.jhi - r.jlo + 1, nk = r.khi - r.klo + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return LSM_OK;
    dim3 blk(64, 4, 1), grd(blocks_for(ni, 64), blocks_for(nj, 4), nk);
    if (mode == LSM_ADVECT_EXACT) weno3_rhs_kernel<false><<<grd, blk, 0, s>>>(a);
    else weno3_rhs_kernel<true><<<grd, blk, 0, s>>>(a);
    return check_hip(hipGetLastError(), "weno3_rhs_kernel");
}

int launch_bcond(const Blk &b, double *phi, hipStream_t s)
{
    BcArgs a{phi, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, make_range(b), {0}, b.big_phi};
    for (int q = 0; q < 6; ++q) a.bt[q] = b.btype[q];
    const long per[3] = {(long)b.e.nj * b.e.nk, (long)b.e.ni * b.e.nk, (long)b.e.ni * b.e.nj};
    for (int ax = 0; ax < 3; ++ax) {
        bcond_phase<<<blocks_for(2 * per[ax], 256), 256, 0, s>>>(a, ax);
        int rc = check_hip(hipGetLastError(), "bcond_phase");
        if (rc) return rc;
    }
    return LSM_OK;
}

static int stage_kc(const Blk &b, const Range &r, int tiles_ij)
{
    // k chunks of 8 planes (measured best at 256^3: 8192 workgroups, the
    // 4-plane halo of a chunk re-read from L2), more if the block is small
    const int nk_int = r.khi - r.klo + 1;
    int chunks = (8192 + tiles_ij - 1) / tiles_ij;
    int kc = (nk_int + chunks - 1) / chunks;
    if (kc < 8) kc = 8;
    if (const char *e = getenv("LSM_KCHUNK")) { int v = atoi(e); if (v > 0) kc = v; }
    return kc;
}

int launch_fold(const Blk &b, const double *U1, const double *U2, const double *U3, double dt,
                double *V1, double *V2, double *V3, int *bad, hipStream_t s)
{
    fold_kernel<<<blocks_for(b.n, 256), 256, 0, s>>>(b.n, b.aj, U1, U2, U3, dt, b.dc, b.de, b.dz,
                                                     V1, V2, V3, bad);
    return check_hip(hipGetLastError(), "fold_kernel");
}

// the three stages of one RK3 step, stage st reading src[st] and writing dst[st]
static int rk3_run(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                   const double *U2, const double *U3, double dt, int mode, const double *const src[3],
                   double *const dst[3], hipStream_t s, bool folded)
{
    if (mode == LSM_ADVECT_UPWIND1) return launch_upwind_rk3(b, phi_n, mask, U1, U2, U3, dt, src, dst, s);
    const Range r = make_range(b);
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const bool fast = mode != LSM_ADVECT_EXACT;

    // interior stage kernel, then the boundary shell kernel, per stage
    int TY = 4;
    if (!folded)
        if (const char *e = getenv("LSM_TILE_Y")) { int v = atoi(e); if (v == 8) TY = 8; }
    bool nt_off = false;
    if (const char *e = getenv("LSM_NT")) nt_off = atoi(e) == 0;
    const int tiles_i = (b.e.ni + TX - 1) / TX;
    const int tiles_j = (b.e.nj + TY - 1) / TY;
    const int kc = stage_kc(b, r, tiles_i * tiles_j);
    const int nk_int = r.khi - r.klo + 1;
    const int tiles_k = nk_int > 0 ? (nk_int + kc - 1) / kc : 0;
    const int ntiles = tiles_i * tiles_j * tiles_k;
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);

    ShellArgs sh{};
    sh.ni = b.e.ni; sh.nj = b.e.nj; sh.nk = b.e.nk; sh.sj = b.e.sj; sh.sk = b.e.sk; sh.r = r;
    for (int q = 0; q < 6; ++q) sh.bt[q] = b.btype[q];
    sh.big = b.big_phi; sh.aj = b.aj; sh.mask = mask; sh.dt = dt;
    shell_counts(r, sh);
    const long nshell = sh.nI + sh.nJ + sh.nK;

    for (int st = 0; st < 3; ++st) {
        StageArgs a{src[st], phi_n, dst[st], U1, U2, U3, b.aj, mask, b.e.ni, b.e.nj, b.e.nk,
                    b.e.sj, b.e.sk, r, C[st][0], C[st][1], C[st][2], dt, b.dc, b.de, b.dz,
                    tiles_i, tiles_j, kc, ntiles, 0};
        if (ntiles > 0) {
            KTimer kt(LSM_KT_ADVECT_STAGE, s);
            dim3 blk(TX, TY, 1);
            if (folded) {
                if (mask) advect_stage<true, true, 4, true, true><<<grid, blk, 0, s>>>(a);
                else advect_stage<true, false, 4, true, true><<<grid, blk, 0, s>>>(a);
            } else if (TY == 8) {
                if (fast && mask) advect_stage<true, true, 8><<<grid, blk, 0, s>>>(a);
                else if (fast) advect_stage<true, false, 8><<<grid, blk, 0, s>>>(a);
                else if (mask) advect_stage<false, true, 8><<<grid, blk, 0, s>>>(a);
                else advect_stage<false, false, 8><<<grid, blk, 0, s>>>(a);
            } else {
                if (fast && mask) advect_stage<true, true, 4><<<grid, blk, 0, s>>>(a);
                else if (fast && nt_off) advect_stage<true, false, 4, false><<<grid, blk, 0, s>>>(a);
                else if (fast) advect_stage<true, false, 4><<<grid, blk, 0, s>>>(a);
                else if (mask) advect_stage<false, true, 4><<<grid, blk, 0, s>>>(a);
                else advect_stage<false, false, 4><<<grid, blk, 0, s>>>(a);
            }
            int rc = check_hip(hipGetLastError(), "advect_stage");
            if (rc) return rc;
        }
        sh.out_ro = dst[st]; sh.out = dst[st]; sh.ps = src[st]; sh.pn = phi_n;
        sh.c1 = C[st][0]; sh.c2 = C[st][1]; sh.c3 = C[st][2];
        if (nshell > 0) {
            KTimer kt(LSM_KT_ADVECT_SHELL, s);
            if (mask) stage_shell<true><<<blocks_for(nshell, 256), 256, 0, s>>>(sh);
            else stage_shell<false><<<blocks_for(nshell, 256), 256, 0, s>>>(sh);
            int rc = check_hip(hipGetLastError(), "stage_shell");
            if (rc) return rc;
        }
    }
    return LSM_OK;
}

int launch_rk3(const Blk &b, double *phi_np1, const double *phi_n, const int32_t *mask,
               const double *U1, const double *U2, const double *U3, double dt, int mode,
               double *scratch, hipStream_t s, bool folded)
{
    const double *src[3] = {phi_n, phi_np1, scratch};
    double *dst[3] = {phi_np1, scratch, phi_np1};
    return rk3_run(b, phi_n, mask, U1, U2, U3, dt, mode, src, dst, s, folded);
}

int launch_rk3_stages(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                      const double *U2, const double *U3, double dt, int mode, double *out1,
                      double *out2, double *out3, hipStream_t s)
{
    const double *src[3] = {phi_n, out1, out2};
    double *dst[3] = {out1, out2, out3};
    return rk3_run(b, phi_n, mask, U1, U2, U3, dt, mode, src, dst, s, false);
}

}  // namespace lsm
