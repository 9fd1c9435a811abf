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
#include "lsm_multi.hpp"

namespace lsm {

namespace {

// NEARC: cached stores next to the physical faces (single blocks; the
// per-block form of a batch launches the variant without them)
template <bool FAST, bool MASK, int TY, bool NT = true, bool FOLD = false, bool NEARC = true>
struct KAdvect {
    using Args = StageArgs;
    static constexpr int kThreads = TX * TY;
    __device__ __forceinline__ static void run(const Args &a, int bid, int nb)
    {
        const int tile = xcd_tile(bid, nb);
        if (tile >= a.ntiles) return;
        advect_tile<FAST, MASK, TY, NT, FOLD, TX, NEARC>(a, tile);
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

template <bool FAST, bool MASK, int TY, bool NT = true, bool FOLD = false, bool NEARC = true>
__global__ __launch_bounds__(TX * TY) void advect_stage(StageArgs a)
{
    KAdvect<FAST, MASK, TY, NT, FOLD, NEARC>::run(a, (int)blockIdx.x, (int)gridDim.x);
}

template <bool FAST, bool MASK, int TY, bool NT, bool FOLD, bool NEARC>
void KAdvect<FAST, MASK, TY, NT, FOLD, NEARC>::one(unsigned g, dim3 b, hipStream_t s, const Args &a)
{
    advect_stage<FAST, MASK, TY, NT, FOLD, NEARC><<<g, b, 0, s>>>(a);
}

template <bool MASK>
struct KShell {
    using Args = ShellArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long t = (long)bid * 256 + threadIdx.x;
        int i, j, k;
        if (!shell_node(a, t, i, j, k)) return;
        const double v = sh_f3<MASK>(a, i, j, k);
        a.out[(long)i + a.sj * j + a.sk * k] = v;
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

template <bool MASK>
__global__ __launch_bounds__(256) void stage_shell(ShellArgs a)
{
    KShell<MASK>::run(a, (int)blockIdx.x, (int)gridDim.x);
}

template <bool MASK>
void KShell<MASK>::one(unsigned g, dim3 b, hipStream_t s, const Args &a)
{
    stage_shell<MASK><<<g, b, 0, s>>>(a);
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

__device__ __forceinline__ void bcond_phase_run(const BcArgs &a, int axis, int bid)
{
    const long t = (long)bid * 256 + threadIdx.x;
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

struct BcPhaseArgs {
    BcArgs a;
    int axis;
};

struct KBcPhase {
    using Args = BcPhaseArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &x, int bid, int) { bcond_phase_run(x.a, x.axis, bid); }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &x);
};

__global__ __launch_bounds__(256) void bcond_phase(BcArgs a, int axis)
{
    bcond_phase_run(a, axis, (int)blockIdx.x);
}

void KBcPhase::one(unsigned g, dim3 b, hipStream_t s, const Args &x) { bcond_phase<<<g, b, 0, s>>>(x.a, x.axis); }

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
    Blank bl;
};

// the WENO3 right-hand side at interior node (i, j, k); the stencils are
// forced one-sided on the physical faces and next to the i and j faces of
// blanking boxes
template <bool FAST>
__device__ __forceinline__ double weno3_node(const RhsArgs &a, const double *p, int i, int j, int k, long l)
{
    const long sj = a.sj, sk = a.sk;
    bool fmb = a.r.fb && i == a.r.ilo, fmf = a.r.ff && i == a.r.ihi;
    bool fml = a.r.fl && j == a.r.jlo, fmr = a.r.fr && j == a.r.jhi;
    const bool fmd = a.r.fd && k == a.r.klo, fmu = a.r.fu && k == a.r.khi;
    if (a.bl.n) a.bl.bias(i, j, fmb, fmf, fml, fmr);
    // forced stencils do not read their far entry: a dummy replaces it, so
    // the reads past a physical face are never made
    const double xm2 = fmb ? 0.0 : p[l - 2], xp2 = fmf ? 0.0 : p[l + 2];
    const double ym2 = fml ? 0.0 : p[l - 2 * sj], yp2 = fmr ? 0.0 : p[l + 2 * sj];
    const double zm2 = fmd ? 0.0 : p[l - 2 * sk], zp2 = fmu ? 0.0 : p[l + 2 * sk];
    const double c = p[l];
    const double Lc = weno3_dir<FAST>(xm2, p[l - 1], c, p[l + 1], xp2, a.U1[l], fmb, fmf, a.dc);
    const double Le = weno3_dir<FAST>(ym2, p[l - sj], c, p[l + sj], yp2, a.U2[l], fml, fmr, a.de);
    const double Lz = weno3_dir<FAST>(zm2, p[l - sk], c, p[l + sk], zp2, a.U3[l], fmd, fmu, a.dz);
    return (Lc + Le) + Lz;
}

// interior nodes with AdvectionNodes != 0 outside the blanking boxes
template <bool FAST>
__global__ __launch_bounds__(256) void weno3_rhs_kernel(RhsArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + a.r.ilo;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + a.r.jlo;
    const int k = blockIdx.z + a.r.klo;
    if (i > a.r.ihi || j > a.r.jhi || k > a.r.khi) return;
    const long l = (long)i + a.sj * j + a.sk * k;
    if (a.mask && a.mask[l] == 0) return;
    if (a.bl.n && a.bl.inside(i, j, k)) return;
    a.rhs[l] = weno3_node<FAST>(a, a.phi, i, j, k, l);
}

// One RK stage of ad_RKTVD3.F90:52-80 on a block with blanking regions,
// unfused: dst = ((c1 phi^n) + (c2 src)) - (((c3 aj) dt) rh(src)) at every
// node with AdvectionNodes > 0 (rh = 0 on the shell and inside the boxes),
// dst = src elsewhere; bcond_lsm follows as its own launches.
struct BlankStageArgs {
    RhsArgs r;
    const double *pn;
    double *dst;
    const double *aj;
    double c1, c2, c3, dt;
};

template <bool FAST>
__global__ __launch_bounds__(256) void blank_stage_kernel(BlankStageArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.r.ni || j >= a.r.nj) return;
    const long l = (long)i + a.r.sj * j + a.r.sk * k;
    const double v = a.r.phi[l];
    if (a.r.mask && !(a.r.mask[l] > 0)) { a.dst[l] = v; return; }
    double rh = 0.0;
    const Range &R = a.r.r;
    if (i >= R.ilo && i <= R.ihi && j >= R.jlo && j <= R.jhi && k >= R.klo && k <= R.khi &&
        !a.r.bl.inside(i, j, k))
        rh = weno3_node<FAST>(a.r, a.r.phi, i, j, k, l);
    a.dst[l] = (a.c1 * a.pn[l] + a.c2 * v) - ((a.c3 * a.aj[l]) * a.dt) * rh;
}

// ---------------------------------------------------------------------------
// the blanking part of bcond_lsm (bcond_lsm.F90:290-562), one box at a time:
// phase 0 zeroes the nodes two and more inside the box, phases 1-4 the i
// faces (ia, ib) and j faces (ja, jb), each face node and the node behind it
// set from the two nodes outside, phase 5 the four i-j corners in the order
// (ia, ja), (ib, ja), (ib, jb), (ia, jb).  k faces are not treated
// (commented out in the reference).
// ---------------------------------------------------------------------------
struct BlankBcArgs {
    double *phi;
    long sj, sk;
    int i0, i1, j0, j1, k0, k1;   // the thread range of the phase (inclusive)
    int ia, ib, ja, jb;           // the box's i and j faces
    int corner[4];                // phase 5: which corners are treated
    double big;
    bool switch_order;            // first order where |phi| at the far node >= big (bcond_lsm)
};

__device__ __forceinline__ void blank_extrap(const BlankBcArgs &a, long lb, long behind, long l1, long l2)
{
    const double v = a.switch_order ? bc_face(1, a.phi[l1], a.phi[l2], 0.0, a.big, false)
                                    : (4.0 / 3.0) * a.phi[l1] + (-1.0 / 3.0) * a.phi[l2];
    a.phi[lb] = v;
    a.phi[behind] = v;
}

__global__ __launch_bounds__(256) void blank_bcond_phase(BlankBcArgs a, int phase)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1 = a.i1 - a.i0 + 1, n2 = a.j1 - a.j0 + 1, n3 = a.k1 - a.k0 + 1;
    if (t >= n1 * n2 * n3) return;
    const int i = a.i0 + (int)(t % n1), j = a.j0 + (int)((t / n1) % n2), k = a.k0 + (int)(t / (n1 * n2));
    const long sj = a.sj, sk = a.sk;
    const long l = (long)i + sj * j + sk * k;
    switch (phase) {
    case 0: a.phi[l] = 0.0; break;
    case 1: blank_extrap(a, l, l + 1, l - 1, l - 2); break;
    case 2: blank_extrap(a, l, l - 1, l + 1, l + 2); break;
    case 3: blank_extrap(a, l, l + sj, l - sj, l - 2 * sj); break;
    case 4: blank_extrap(a, l, l - sj, l + sj, l + 2 * sj); break;
    default: {
        // one thread per k plane, the corners in the reference's order
        const long k0 = sk * k;
        if (a.corner[0]) {
            const long c = a.ia + sj * a.ja + k0;
            blank_extrap(a, c, c + 1 + sj, c - 1 - sj, c - 2 - 2 * sj);
        }
        if (a.corner[1]) {
            const long c = a.ib + sj * a.ja + k0;
            blank_extrap(a, c, c - 1 + sj, c + 1 - sj, c + 2 - 2 * sj);
        }
        if (a.corner[2]) {
            const long c = a.ib + sj * a.jb + k0;
            blank_extrap(a, c, c - 1 - sj, c + 1 + sj, c + 2 + 2 * sj);
        }
        if (a.corner[3]) {
            const long c = a.ia + sj * a.jb + k0;
            blank_extrap(a, c, c + 1 - sj, c - 1 + sj, c - 2 + 2 * sj);
        }
    }
    }
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

struct KContra {
    using Args = ContraArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int);
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void contra_kernel(ContraArgs a) { KContra::run(a, (int)blockIdx.x, (int)gridDim.x); }

void KContra::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { contra_kernel<<<g, b, 0, s>>>(a); }

__device__ __forceinline__ void KContra::run(const Args &a, int bid, int)
{
    const long l = (long)bid * 256 + threadIdx.x;
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

struct CopyArgs {
    const double *src;
    double *dst;
    long n;
};

struct KCopy {
    using Args = CopyArgs;
    static constexpr int kThreads = 256;
    __device__ __forceinline__ static void run(const Args &a, int bid, int)
    {
        const long l = (long)bid * 256 + threadIdx.x;
        if (l < a.n) a.dst[l] = a.src[l];
    }
    static void one(unsigned g, dim3 b, hipStream_t s, const Args &a);
};

__global__ __launch_bounds__(256) void copy_kernel(const double *__restrict__ s, double *__restrict__ d, long n)
{
    KCopy::run(CopyArgs{s, d, n}, (int)blockIdx.x, (int)gridDim.x);
}

void KCopy::one(unsigned g, dim3 b, hipStream_t s, const Args &a) { copy_kernel<<<g, b, 0, s>>>(a.src, a.dst, a.n); }

inline unsigned blocks_for(long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// one block's stage tiles and shell with the single-block kernels (the
// per-block form of the batched step)
int launch_stage_block(const StageArgs &a, int ty, bool fast, hipStream_t s)
{
    const unsigned grid = (unsigned)((a.ntiles + 7) / 8 * 8);
    const dim3 blk(TX, ty, 1);
    if (ty == 8) {
        if (fast) advect_stage<true, false, 8, true, false, false><<<grid, blk, 0, s>>>(a);
        else advect_stage<false, false, 8, true, false, false><<<grid, blk, 0, s>>>(a);
    } else {
        if (fast) advect_stage<true, false, 4, true, false, false><<<grid, blk, 0, s>>>(a);
        else advect_stage<false, false, 4, true, false, false><<<grid, blk, 0, s>>>(a);
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
    return klaunch<KCopy>(blocks_for(n, 256), dim3(256), s, CopyArgs{src, dst, n}, "copy_kernel");
}

int launch_contra(const Blk &b, const double *u, const double *v, const double *w, long stride,
                  double *U1, double *U2, double *U3, hipStream_t s)
{
    ContraArgs a{u, v, w, stride, b.csi0, b.csi1, b.csi2, b.eta0, b.eta1, b.eta2,
                 b.zet0, b.zet1, b.zet2, b.aj, U1, U2, U3, b.n};
    return klaunch<KContra>(blocks_for(b.n, 256), dim3(256), s, a, "contra_kernel");
}

static RhsArgs rhs_args(const Blk &b, const double *phi, const double *U1, const double *U2, const double *U3,
                        const int32_t *mask, double *rhs)
{
    return RhsArgs{phi, U1, U2, U3, mask, rhs, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, make_range(b),
                   b.dc, b.de, b.dz, b.bl};
}

int launch_weno3_rhs(const Blk &b, const double *phi, const double *U1, const double *U2,
                     const double *U3, const int32_t *mask, double *rhs, int mode, hipStream_t s)
{
    const RhsArgs a = rhs_args(b, phi, U1, U2, U3, mask, rhs);
    const Range &r = a.r;
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
        int rc = klaunch<KBcPhase>(blocks_for(2 * per[ax], 256), dim3(256), s, BcPhaseArgs{a, ax}, "bcond_phase");
        if (rc) return rc;
    }
    if (b.bl.n && recording()) return fail(LSM_ERR_UNSUPPORTED, "blanking regions cannot be recorded");
    return launch_bcond_blanking(b, phi, s);
}

int launch_bcond_blanking(const Blk &b, double *phi, hipStream_t s)
{
    return launch_blank_extrap(b, phi, 0, true, true, false, s);
}

// The phases of one box after another; a phase whose guard fails or whose
// range is empty is not launched.  m: the first owned index the guards and
// the face loops count from (bcond_lsm's i_mysta = il: 0, calc_h's
// il + 1: 1), so a face is treated where it lies more than m + 1 nodes
// inside the block, and its loops are clamped to [m, n - 1 - m].
int launch_blank_extrap(const Blk &b, double *phi, int m, bool zero, bool switch_order, bool full_k,
                        hipStream_t s)
{
    const int ni = b.e.ni, nj = b.e.nj, nk = b.e.nk;
    const int mei = ni - 1 - m, mej = nj - 1 - m;
    for (int q = 0; q < b.bl.n; ++q) {
        const int *B = b.bl.b[q];
        const int ia = B[0], ib = B[1], ja = B[2], jb = B[3], ka = B[4], kb = B[5];
        // bcond_lsm.F90:324, 358, 386, 410 (calc_h.F90:215, 241, 261, 280): blktype 0 and
        // two nodes between the face and the first owned node
        const bool fi0 = B[6] == 0 && ia > m + 1, fi1 = B[7] == 0 && ib < mei - 1;
        const bool fj0 = B[8] == 0 && ja > m + 1, fj1 = B[9] == 0 && jb < mej - 1;
        const int k0 = full_k ? 0 : max(0, ka), k1 = full_k ? nk - 1 : min(nk - 1, kb);
        BlankBcArgs a{};
        a.phi = phi; a.sj = b.e.sj; a.sk = b.e.sk; a.big = b.big_phi; a.switch_order = switch_order;
        a.ia = ia; a.ib = ib; a.ja = ja; a.jb = jb;
        a.corner[0] = fi0 && fj0; a.corner[1] = fi1 && fj0; a.corner[2] = fi1 && fj1; a.corner[3] = fi0 && fj1;
        auto run = [&](int phase, bool on, int i0, int i1, int j0, int j1, int kk0, int kk1) -> int {
            if (!on || i0 > i1 || j0 > j1 || kk0 > kk1) return LSM_OK;
            a.i0 = i0; a.i1 = i1; a.j0 = j0; a.j1 = j1; a.k0 = kk0; a.k1 = kk1;
            const long n = (long)(i1 - i0 + 1) * (j1 - j0 + 1) * (kk1 - kk0 + 1);
            blank_bcond_phase<<<blocks_for(n, 256), 256, 0, s>>>(a, phase);
            return check_hip(hipGetLastError(), "blank_bcond_phase");
        };
        const int jlo = max(m, ja), jhi = min(mej, jb), ilo = max(m, ia), ihi = min(mei, ib);
        int rc;
        if ((rc = run(0, zero, max(0, ia + 2), min(ni - 1, ib - 2), max(0, ja + 2), min(nj - 1, jb - 2),
                      max(0, ka + 2), min(nk - 1, kb - 2)))) return rc;
        if ((rc = run(1, fi0, ia, ia, jlo, jhi, k0, k1))) return rc;
        if ((rc = run(2, fi1, ib, ib, jlo, jhi, k0, k1))) return rc;
        if ((rc = run(3, fj0, ilo, ihi, ja, ja, k0, k1))) return rc;
        if ((rc = run(4, fj1, ilo, ihi, jb, jb, k0, k1))) return rc;
        const bool anyc = a.corner[0] || a.corner[1] || a.corner[2] || a.corner[3];
        if ((rc = run(5, anyc, 0, 0, 0, 0, k0, k1))) return rc;
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
    if (recording() && (mode == LSM_ADVECT_UPWIND1 || b.bl.n))
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 and blanking regions cannot be recorded");
    if (mode == LSM_ADVECT_UPWIND1) return launch_upwind_rk3(b, phi_n, mask, U1, U2, U3, dt, src, dst, s);
    const Range r = make_range(b);
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const bool fast = mode != LSM_ADVECT_EXACT;
    if (b.bl.n) {
        if (folded) return fail(LSM_ERR_UNSUPPORTED, "blanking regions: the folded step is not implemented");
        if (b.e.nk > 65535) return fail(LSM_ERR_ARG, "nk > 65535");
        for (int st = 0; st < 3; ++st) {
            BlankStageArgs a{rhs_args(b, src[st], U1, U2, U3, mask, nullptr), phi_n, dst[st], b.aj,
                             C[st][0], C[st][1], C[st][2], dt};
            const dim3 grd(blocks_for(b.e.ni, 64), blocks_for(b.e.nj, 4), b.e.nk);
            {
                KTimer kt(LSM_KT_ADVECT_STAGE, s);
                if (fast) blank_stage_kernel<true><<<grd, dim3(64, 4, 1), 0, s>>>(a);
                else blank_stage_kernel<false><<<grd, dim3(64, 4, 1), 0, s>>>(a);
            }
            int rc = check_hip(hipGetLastError(), "blank_stage_kernel");
            if (rc) return rc;
            if ((rc = launch_bcond(b, dst[st], s))) return rc;
        }
        return LSM_OK;
    }

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
                    tiles_i, tiles_j, kc, ntiles, 0, 0};
        if (ntiles > 0) {
            KTimer kt(LSM_KT_ADVECT_STAGE, s);
            dim3 blk(TX, TY, 1);
            int rc;
            if (folded) {
                if (mask) rc = klaunch<KAdvect<true, true, 4, true, true>>(grid, blk, s, a, "advect_stage");
                else rc = klaunch<KAdvect<true, false, 4, true, true>>(grid, blk, s, a, "advect_stage");
            } else if (TY == 8) {
                if (fast && mask) rc = klaunch<KAdvect<true, true, 8>>(grid, blk, s, a, "advect_stage");
                else if (fast) rc = klaunch<KAdvect<true, false, 8>>(grid, blk, s, a, "advect_stage");
                else if (mask) rc = klaunch<KAdvect<false, true, 8>>(grid, blk, s, a, "advect_stage");
                else rc = klaunch<KAdvect<false, false, 8>>(grid, blk, s, a, "advect_stage");
            } else {
                if (fast && mask) rc = klaunch<KAdvect<true, true, 4>>(grid, blk, s, a, "advect_stage");
                else if (fast && nt_off) rc = klaunch<KAdvect<true, false, 4, false>>(grid, blk, s, a, "advect_stage");
                else if (fast) rc = klaunch<KAdvect<true, false, 4>>(grid, blk, s, a, "advect_stage");
                else if (mask) rc = klaunch<KAdvect<false, true, 4>>(grid, blk, s, a, "advect_stage");
                else rc = klaunch<KAdvect<false, false, 4>>(grid, blk, s, a, "advect_stage");
            }
            if (rc) return rc;
        }
        sh.out_ro = dst[st]; sh.out = dst[st]; sh.ps = src[st]; sh.pn = phi_n;
        sh.c1 = C[st][0]; sh.c2 = C[st][1]; sh.c3 = C[st][2];
        if (nshell > 0) {
            KTimer kt(LSM_KT_ADVECT_SHELL, s);
            int rc = mask ? klaunch<KShell<true>>(blocks_for(nshell, 256), dim3(256), s, sh, "stage_shell")
                          : klaunch<KShell<false>>(blocks_for(nshell, 256), dim3(256), s, sh, "stage_shell");
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
