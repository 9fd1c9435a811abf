// The interior tile of one fused WENO3 + SSP-RK3 stage and the boundary shell
// of a stage (ad_RKTVD3.F90:62-84, calc_RH_AD_WENO3.F90:86-403,
// bcond_lsm.F90:26-287), shared by the single-block kernels and the batched
// multi-block kernels.  A block's arrays may carry ghost layers: `Range` gives
// its owned nodes (ia..ib, ...), the nodes whose right-hand side is computed
// (ilo..ihi, ...: owned nodes not on a physical face) and which faces are
// physical.
#pragma once

#include "lsm_internal.hpp"

namespace lsm {

constexpr int TX = 64;    // tile width in i (one wavefront per j row)
constexpr int HALO = 2;

struct Range {
    int ilo, ihi, jlo, jhi, klo, khi;   // interior nodes (RHS nodes)
    int ia, ib, ja, jb, ka, kb;         // owned nodes
    bool fb, ff, fl, fr, fd, fu;        // physical faces
};

inline Range make_range(const Blk &b)
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
    int jb0;                // j of the first tile row
    int ib0;                // i of the first tile column (a 128-byte line boundary of the rows)
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

// one tile of the stage: W (i) x TY (j) columns of interior nodes marching
// up a chunk of kc planes.  W = 64: one wavefront per j row (blockDim 64 x
// TY); W = 32: each wavefront covers two half rows (blockDim 64 x TY/2), for
// blocks whose rows are not much longer than a multiple of 32
// NEARC: cached stores for the two node layers next to a physical face (single
// blocks, whose boundary shell reads them right after the stage; in the
// batched forms they displaced the halo lines the tiles reuse)
template <bool FAST, bool MASK, int TY, bool NT = true, bool FOLD = false, int W = TX, bool NEARC = false>
__device__ __forceinline__ void advect_tile(const StageArgs &a, int tile)
{
    const int ti = tile % a.tiles_i;
    const int t2 = tile / a.tiles_i;
    const int tj = t2 % a.tiles_j;
    const int tk = t2 / a.tiles_j;
    int tx, ty;
    if constexpr (W == TX) { tx = threadIdx.x; ty = threadIdx.y; }
    else { tx = threadIdx.x & (W - 1); ty = threadIdx.y * (TX / W) + (threadIdx.x / W); }
    const int i = a.ib0 + ti * W + tx, j = a.jb0 + tj * TY + ty;
    const int kb = a.r.klo + tk * a.kc;
    const int ke = min(kb + a.kc - 1, a.r.khi);

    __shared__ double sm[2][TY + 2 * HALO][W + 2 * HALO];

    // Loads are clamped into the array instead of guarded: a clamped value is
    // only ever read by a thread that does not compute (outside the interior)
    // or by a stencil slot the boundary forcing replaces with a dummy, and
    // unguarded loads issue back to back without branches.
    const int ic = min(i, a.ni - 1), jc = min(j, a.nj - 1);
    const long col = (long)ic + a.sj * jc;
    // waves 0,1 load the two halo rows below the tile, waves 2,3 the two above
    const bool has_jh = ty < 4;
    const int rr = (ty < 2) ? ty : ty - 2 + TY + HALO;
    const int jh = min(max(a.jb0 + tj * TY - HALO + rr, 0), a.nj - 1);
    const long col_jh = (long)ic + a.sj * jh;
    const int lin = ty * W + tx;
    const bool has_ih = lin < 4 * TY;                       // 16 i-halo cells
    const int hr = (lin >> 2) % TY, hc4 = lin & 3;
    const int hcol = (hc4 < 2) ? hc4 : hc4 + W;
    const int hi_i = min(max(a.ib0 + ti * W - HALO + hcol, 0), a.ni - 1);
    const int hi_j = min(a.jb0 + tj * TY + hr, a.nj - 1);
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
                const double (*p)[W + 2 * HALO] = sm[buf];
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
            // the two layers next to a physical face are read back by the
            // boundary shell right after the stage: cached stores there
            const bool nearf = NEARC && ((a.r.fb && i <= a.r.ilo + 1) || (a.r.ff && i >= a.r.ihi - 1) ||
                               (a.r.fl && j <= a.r.jlo + 1) || (a.r.fr && j >= a.r.jhi - 1) ||
                               (a.r.fd && k <= a.r.klo + 1) || (a.r.fu && k >= a.r.khi - 1));
            if (nearf) a.out[col + a.sk * (long)k] = v;
            else stnt(a.out + col + a.sk * (long)k, v);
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

// decode shell node t: i faces own all their owned (j,k); j faces own the
// owned i not on a physical i face; k faces own the owned (i,j) not on a
// physical i or j face
__device__ __forceinline__ bool shell_node(const ShellArgs &a, long t, int &i, int &j, int &k)
{
    const int ilo2 = a.r.ia + (a.r.fb ? 1 : 0);
    const int jlo2 = a.r.ja + (a.r.fl ? 1 : 0);
    const int noj = a.r.jb - a.r.ja + 1, nok = a.r.kb - a.r.ka + 1;
    if (t < a.nI) {
        const long per = (long)noj * nok;
        const int f = (int)(t / per);
        const long rem = t % per;
        j = a.r.ja + (int)(rem % noj);
        k = a.r.ka + (int)(rem / noj);
        i = (f == 0 && a.r.fb) ? a.r.ia : a.r.ib;
        return true;
    }
    t -= a.nI;
    if (t < a.nJ) {
        const long per = (long)a.cnt_i * nok;
        const int f = (int)(t / per);
        const long rem = t % per;
        i = ilo2 + (int)(rem % a.cnt_i);
        k = a.r.ka + (int)(rem / a.cnt_i);
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

// shell launch parameters of one block: node counts per face group
inline void shell_counts(const Range &r, ShellArgs &sh)
{
    const int nfi = (r.fb ? 1 : 0) + (r.ff ? 1 : 0);
    const int nfj = (r.fl ? 1 : 0) + (r.fr ? 1 : 0);
    const int nfk = (r.fd ? 1 : 0) + (r.fu ? 1 : 0);
    const int noi = r.ib - r.ia + 1, noj = r.jb - r.ja + 1, nok = r.kb - r.ka + 1;
    sh.cnt_i = noi - nfi;
    sh.cnt_j = noj - nfj;
    sh.nI = (long)nfi * noj * nok;
    sh.nJ = (long)nfj * sh.cnt_i * nok;
    sh.nK = (long)nfk * sh.cnt_i * sh.cnt_j;
}

}  // namespace lsm
