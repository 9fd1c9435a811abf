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

namespace lsm {

namespace {

constexpr int TX = 64;    // tile width in i (one wavefront per j row)
constexpr int HALO = 2;

struct Range {
    int ilo, ihi, jlo, jhi, klo, khi;   // interior nodes (RHS nodes)
    int ia, ib, ja, jb, ka, kb;         // owned nodes
    bool fb, ff, fl, fr, fd, fu;        // physical faces
};

Range make_range(const Blk &b)
{
    Range r;
    const int f = b.e.faces;
    r.fb = f & LSM_FACE_I1; r.ff = f & LSM_FACE_IM;
    r.fl = f & LSM_FACE_J1; r.fr = f & LSM_FACE_JM;
    r.fd = f & LSM_FACE_K1; r.fu = f & LSM_FACE_KM;
    r.ia = 0; r.ib = b.e.ni - 1;
    r.ja = 0; r.jb = b.e.nj - 1;
    r.ka = 0; r.kb = b.e.nk - 1;
    r.ilo = r.ia + (r.fb ? 1 : 0); r.ihi = r.ib - (r.ff ? 1 : 0);
    r.jlo = r.ja + (r.fl ? 1 : 0); r.jhi = r.jb - (r.fr ? 1 : 0);
    r.klo = r.ka + (r.fd ? 1 : 0); r.khi = r.kb - (r.fu ? 1 : 0);
    return r;
}

// workgroup -> tile, keeping consecutive tiles on one XCD: the dispatcher
// deals workgroups round-robin over the 8 XCDs, so workgroup b runs on XCD
// b % 8; giving each XCD a contiguous run of tiles keeps the j/k halo reuse
// inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

struct StageArgs {
    const double *ps, *pn;
    double *out;
    const double *U1, *U2, *U3, *aj;
    const int32_t *mask;
    int ni, nj, nk;
    long sj, sk;
    Range r;
    double c1, c2, c3, dt, dc, de, dz;
    int tiles_i, tiles_j, kc, ntiles;
};

// streamed operands are used once per stage: non-temporal loads/stores keep
// them from evicting the phi planes whose halos neighbouring tiles re-read
// (LSM_NT=0 turns this off)
template <bool NT>
__device__ __forceinline__ double ldnt_(const double *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT>
__device__ __forceinline__ void stnt_(double *p, double v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool FAST, bool MASK, int TY, bool NT = true, bool FOLD = false>
__global__ __launch_bounds__(TX * TY) void advect_stage(StageArgs a)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= a.ntiles) return;
    const int ti = tile % a.tiles_i;
    const int t2 = tile / a.tiles_i;
    const int tj = t2 % a.tiles_j;
    const int tk = t2 / a.tiles_j;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = ti * TX + tx, j = tj * TY + ty;
    const int kb = a.r.klo + tk * a.kc;
    const int ke = min(kb + a.kc - 1, a.r.khi);

    __shared__ double sm[2][TY + 2 * HALO][TX + 2 * HALO];

    // Loads are clamped into the array instead of guarded: a clamped value is
    // only ever read by a thread that does not compute (outside the interior)
    // or by a stencil slot the boundary forcing replaces with a dummy, and
    // unguarded loads issue back to back without branches.
    const int ic = min(i, a.ni - 1), jc = min(j, a.nj - 1);
    const long col = (long)ic + a.sj * jc;
    // waves 0,1 load the two halo rows below the tile, waves 2,3 the two above
    const bool has_jh = ty < 4;
    const int rr = (ty < 2) ? ty : ty - 2 + TY + HALO;
    const int jh = min(max(tj * TY - HALO + rr, 0), a.nj - 1);
    const long col_jh = (long)ic + a.sj * jh;
    const int lin = ty * TX + tx;
    const bool has_ih = lin < 4 * TY;                       // 16 i-halo cells
    const int hr = (lin >> 2) % TY, hc4 = lin & 3;
    const int hcol = (hc4 < 2) ? hc4 : hc4 + TX;
    const int hi_i = min(max(ti * TX - HALO + hcol, 0), a.ni - 1);
    const int hi_j = min(tj * TY + hr, a.nj - 1);
    const long col_ih = (long)hi_i + a.sj * hi_j;
    auto kcl = [&](int kk) -> long { return a.sk * (long)min(max(kk, 0), a.nk - 1); };

    const bool act = i >= a.r.ilo && i <= a.r.ihi && j >= a.r.jlo && j <= a.r.jhi;
    const bool fmb = a.r.fb && i == a.r.ilo, fmf = a.r.ff && i == a.r.ihi;
    const bool fml = a.r.fl && j == a.r.jlo, fmr = a.r.fr && j == a.r.jhi;
    const double *__restrict__ ps = a.ps;
    auto ldnt = [](const double *p) { return ldnt_<NT>(p); };
    auto stnt = [](double *p, double v) { stnt_<NT>(p, v); };

    // register queue phi(k-2 .. k+1) of this column; next-plane operands
    double q0 = ps[col + kcl(kb - 2)], q1 = ps[col + kcl(kb - 1)];
    double q2 = ps[col + kcl(kb)], q3 = ps[col + kcl(kb + 1)];
    double nq4 = ps[col + kcl(kb + 2)];
    double njh = has_jh ? ps[col_jh + kcl(kb)] : 0.0;
    double nih = has_ih ? ps[col_ih + kcl(kb)] : 0.0;
    long l = col + a.sk * (long)kb;
    double nu1 = ldnt(a.U1 + l), nu2 = ldnt(a.U2 + l), nu3 = ldnt(a.U3 + l),
           naj = FOLD ? 0.0 : ldnt(a.aj + l), npn = ldnt(a.pn + l);
    int nmk = MASK ? a.mask[l] : 1;

    int buf = 0;
    for (int k = kb; k <= ke; ++k) {
        const double q4 = nq4;
        const double u1 = nu1, u2 = nu2, u3 = nu3, ajv = naj, pnv = npn;
        const int mk = nmk;
        sm[buf][ty + HALO][tx + HALO] = q2;
        if (has_jh) sm[buf][rr][tx + HALO] = njh;
        if (has_ih) sm[buf][hr + HALO][hcol] = nih;
        __syncthreads();
        // operands of the next plane, in flight while this plane computes
        if (k < ke) {
            const long kn = kcl(k + 1);
            nq4 = ps[col + kcl(k + 3)];
            if (has_jh) njh = ps[col_jh + kn];
            if (has_ih) nih = ps[col_ih + kn];
            const long ln = col + kn;
            nu1 = ldnt(a.U1 + ln); nu2 = ldnt(a.U2 + ln); nu3 = ldnt(a.U3 + ln);
            if (!FOLD) naj = ldnt(a.aj + ln);
            npn = ldnt(a.pn + ln);
            if (MASK) nmk = a.mask[ln];
        }
        if (act) {
            double v = q2;
            if (mk > 0) {
                const double (*p)[TX + 2 * HALO] = sm[buf];
                const int cx = tx + HALO, cy = ty + HALO;
                // folded operator: u1..u3 are aj*dt*d*ucn, so the metric scale is 1
                const double sc = FOLD ? 1.0 : a.dc, se = FOLD ? 1.0 : a.de, sz = FOLD ? 1.0 : a.dz;
                const double Lc = weno3_dir<FAST>(p[cy][cx - 2], p[cy][cx - 1], q2, p[cy][cx + 1],
                                                  p[cy][cx + 2], u1, fmb, fmf, sc);
                const double Le = weno3_dir<FAST>(p[cy - 2][cx], p[cy - 1][cx], q2, p[cy + 1][cx],
                                                  p[cy + 2][cx], u2, fml, fmr, se);
                const bool fmd = a.r.fd && k == a.r.klo, fmu = a.r.fu && k == a.r.khi;
                const double Lz = weno3_dir<FAST>(q0, q1, q2, q3, q4, u3, fmd, fmu, sz);
                const double rh = (Lc + Le) + Lz;
                if (FOLD) v = (a.c1 * pnv + a.c2 * q2) - a.c3 * rh;
                else v = (a.c1 * pnv + a.c2 * q2) - ((a.c3 * ajv) * a.dt) * rh;
            }
            stnt(a.out + col + a.sk * (long)k, v);
        }
        q0 = q1; q1 = q2; q2 = q3; q3 = q4;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------
// boundary shell of one RK stage: RK combination with rh = 0 on physical
// boundary nodes (ad_RKTVD3.F90:62-81) followed by bcond_lsm (ad_RKTVD3.F90:83)
// evaluated per node.  Node values after the i faces, j faces and k faces
// (f1, f2, f3) are expanded recursively from f0 = the stage result before
// boundary conditions, so every thread only reads arrays no thread writes.
// ---------------------------------------------------------------------------
struct ShellArgs {
    const double *out_ro;   // interior results of the stage (read)
    double *out;            // the same array (boundary nodes written)
    const double *ps, *pn, *aj;
    const int32_t *mask;
    int ni, nj, nk;
    long sj, sk;
    Range r;
    int bt[6];
    double big, c1, c2, c3, dt;
    long nI, nJ, nK;        // shell nodes owned by the i, j and k faces
    int cnt_i, cnt_j;       // owned i / j nodes not on a physical i / j face
};

template <bool MASK>
__device__ __forceinline__ double sh_f0(const ShellArgs &a, int i, int j, int k)
{
    const long l = (long)i + a.sj * j + a.sk * k;
    if (i >= a.r.ilo && i <= a.r.ihi && j >= a.r.jlo && j <= a.r.jhi && k >= a.r.klo && k <= a.r.khi)
        return a.out_ro[l];
    if (MASK && a.mask[l] <= 0) return a.ps[l];
    return (a.c1 * a.pn[l] + a.c2 * a.ps[l]) - ((a.c3 * a.aj[l]) * a.dt) * 0.0;
}

template <bool MASK>
__device__ __forceinline__ double sh_f1(const ShellArgs &a, int i, int j, int k)
{
    if (a.r.fb && i == a.r.ia) {
        const int bt = a.bt[0];
        if (bt == 2) return sh_f0<MASK>(a, i, j, k);
        return bc_face(bt, sh_f0<MASK>(a, i + 1, j, k), sh_f0<MASK>(a, i + 2, j, k), 0.0, a.big, false);
    }
    if (a.r.ff && i == a.r.ib) {
        return bc_face(a.bt[1], sh_f0<MASK>(a, i - 1, j, k), sh_f0<MASK>(a, i - 2, j, k), 0.0, a.big, true);
    }
    return sh_f0<MASK>(a, i, j, k);
}

template <bool MASK>
__device__ __forceinline__ double sh_f2(const ShellArgs &a, int i, int j, int k)
{
    if (a.r.fl && j == a.r.ja) {
        const int bt = a.bt[2];
        if (bt == 2) return sh_f1<MASK>(a, i, j, k);
        return bc_face(bt, sh_f1<MASK>(a, i, j + 1, k), sh_f1<MASK>(a, i, j + 2, k), 0.0, a.big, false);
    }
    if (a.r.fr && j == a.r.jb) {
        const int bt = a.bt[3];
        if (bt == 2) return sh_f1<MASK>(a, i, j, k);
        return bc_face(bt, sh_f1<MASK>(a, i, j - 1, k), sh_f1<MASK>(a, i, j - 2, k), 0.0, a.big, false);
    }
    return sh_f1<MASK>(a, i, j, k);
}

template <bool MASK>
__device__ __forceinline__ double sh_f3(const ShellArgs &a, int i, int j, int k)
{
    if (a.r.fd && k == a.r.ka) {
        const int bt = a.bt[4];
        if (bt == 2) return sh_f2<MASK>(a, i, j, k);
        return bc_face(bt, sh_f2<MASK>(a, i, j, k + 1), sh_f2<MASK>(a, i, j, k + 2), 0.0, a.big, false);
    }
    if (a.r.fu && k == a.r.kb) {
        const int bt = a.bt[5];
        if (bt == 2) return sh_f2<MASK>(a, i, j, k);
        return bc_face(bt, sh_f2<MASK>(a, i, j, k - 1), sh_f2<MASK>(a, i, j, k - 2), 0.0, a.big, false);
    }
    return sh_f2<MASK>(a, i, j, k);
}

// decode shell node t: i faces own all their (j,k); j faces own the i not on
// a physical i face; k faces own the (i,j) not on a physical i or j face
__device__ __forceinline__ bool shell_node(const ShellArgs &a, long t, int &i, int &j, int &k)
{
    const int ilo2 = a.r.ia + (a.r.fb ? 1 : 0);
    const int jlo2 = a.r.ja + (a.r.fl ? 1 : 0);
    if (t < a.nI) {
        const long per = (long)a.nj * a.nk;
        const int f = (int)(t / per);
        const long rem = t % per;
        j = (int)(rem % a.nj);
        k = (int)(rem / a.nj);
        i = (f == 0 && a.r.fb) ? a.r.ia : a.r.ib;
        return true;
    }
    t -= a.nI;
    if (t < a.nJ) {
        const long per = (long)a.cnt_i * a.nk;
        const int f = (int)(t / per);
        const long rem = t % per;
        i = ilo2 + (int)(rem % a.cnt_i);
        k = (int)(rem / a.cnt_i);
        j = (f == 0 && a.r.fl) ? a.r.ja : a.r.jb;
        return true;
    }
    t -= a.nJ;
    if (t < a.nK) {
        const long per = (long)a.cnt_i * a.cnt_j;
        const int f = (int)(t / per);
        const long rem = t % per;
        i = ilo2 + (int)(rem % a.cnt_i);
        j = jlo2 + (int)(rem / a.cnt_i);
        k = (f == 0 && a.r.fd) ? a.r.ka : a.r.kb;
        return true;
    }
    return false;
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
    const int nfi = (r.fb ? 1 : 0) + (r.ff ? 1 : 0);
    const int nfj = (r.fl ? 1 : 0) + (r.fr ? 1 : 0);
    const int nfk = (r.fd ? 1 : 0) + (r.fu ? 1 : 0);
    sh.cnt_i = b.e.ni - nfi;
    sh.cnt_j = b.e.nj - nfj;
    sh.nI = (long)nfi * b.e.nj * b.e.nk;
    sh.nJ = (long)nfj * sh.cnt_i * b.e.nk;
    sh.nK = (long)nfk * sh.cnt_i * sh.cnt_j;
    const long nshell = sh.nI + sh.nJ + sh.nK;

    for (int st = 0; st < 3; ++st) {
        StageArgs a{src[st], phi_n, dst[st], U1, U2, U3, b.aj, mask, b.e.ni, b.e.nj, b.e.nk,
                    b.e.sj, b.e.sk, r, C[st][0], C[st][1], C[st][2], dt, b.dc, b.de, b.dz,
                    tiles_i, tiles_j, kc, ntiles};
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
