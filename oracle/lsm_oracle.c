/*
 * lsm_oracle.c -- CPU restatement of the reference level-set transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Every routine restates one reference routine, in the same
 * floating-point operation order (compile with -ffp-contract=off, no
 * fast-math), and cites the reference file:line it follows.  Its results are
 * pinned against golden vectors written by the compiled reference Fortran
 * (tests/golden) and against closed forms (tests/test_oracle_*.py).
 *
 * Layout: one block, every face a physical boundary (one MPI rank in the
 * reference: igp = jgp = kgp = 0), arrays i-fastest with 0-based indices,
 * node (i,j,k) of the reference = (i-1, j-1, k-1) here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t ni, nj, nk;
    int32_t btype[6];          /* btype_lsm(1:6): i1, im, j1, jm, k1, km */
    double dc, de, dz;
    double big_phi;
    const double *csi[3], *eta[3], *zet[3];   /* grad xi, grad eta, grad zeta */
    const double *aj;
    const double *x, *y, *z;
    /* blanking regions (global_app.F90:30-43; one rank: li_blk = le_blk):
       nblk boxes, each (ia, ib, ja, jb, ka, kb) as 0-based node indices,
       then blktype(1:4) of the box */
    int32_t nblk;
    int32_t blank[4][10];
} OBlock;

typedef struct {
    double epslsm;
    double eps_reinit;          /* epsReinitialisation */
    double rn_timestep;         /* RNtimestep */
    int32_t rn_itermax;         /* RNitermax */
    int32_t order_bc;           /* OrderReinitialisationBoundaries */
    int32_t eno_bc;             /* ENOBCReinitialisation */
    int32_t lambda_method;      /* sussman_correction_method (Godunov path): 0 point, 1 filtered */
} OReinit;

#define EPS_SIMS 2.220446049250313e-16   /* epsilon(1._rdf), precision.F90:62 */
#define PI_REF 3.141592653589793         /* precision.F90:51 */

static inline size_t IX(const OBlock *b, int i, int j, int k)
{
    return (size_t)i + (size_t)b->ni * ((size_t)j + (size_t)b->nj * (size_t)k);
}

static inline double fsign1(double x) { return copysign(1.0, x); }

/* ------------------------------------------------------------------------- */
/* scalar helpers                                                            */
/* ------------------------------------------------------------------------- */

/* GetWENO3Reconstruction, AdvectionMethods.F90:195-245 */
double oracle_weno3(double f0, double f1, double f2, double f3, double ws)
{
    const double d0 = 2.0 / 3.0, d1 = 1.0 / 3.0, e = 0.000001;
    double fL, fC, fR;
    if (ws > 0.0) { fL = f0; fC = f1; fR = f2; }
    else          { fL = f3; fC = f2; fR = f1; }
    double b0 = (fC - fR) * (fC - fR);
    double b1 = (fL - fC) * (fL - fC);
    double a0 = d0 / ((b0 + e) * (b0 + e));
    double a1 = d1 / ((b1 + e) * (b1 + e));
    double w0 = a0 / (a0 + a1);
    double w1 = a1 / (a0 + a1);
    double u0 = 0.5 * fC + 0.5 * fR;
    double u1 = (-0.5) * fL + 1.5 * fC;
    return u0 * w0 + u1 * w1;
}

/* GetENO2Reconstruction, AdvectionMethods.F90:64-191 */
double oracle_eno2(double f0, double fLL, double fL, double fC, double fR, double fRR,
                   int left_out, int right_out)
{
    double a, b, Mf, dp, dm, s;
    if (left_out) {
        a = (fR - 2.0 * fC) + fL;
        b = (fRR - 2.0 * fR) + fC;
        Mf = a; if (fabs(b) < fabs(a)) Mf = b;
        dp = (fR - fC) - 0.5 * Mf;
        a = (fR - 2.0 * fC) + fL;
        Mf = a;
        dm = (fC - fL) + 0.5 * Mf;
    } else if (right_out) {
        a = (fR - 2.0 * fC) + fL;
        Mf = a;
        dp = (fR - fC) - 0.5 * Mf;
        a = (fR - 2.0 * fC) + fL;
        b = (fC - 2.0 * fL) + fLL;
        Mf = a; if (fabs(b) < fabs(a)) Mf = b;
        dm = (fC - fL) + 0.5 * Mf;
    } else {
        a = (fR - 2.0 * fC) + fL;
        b = (fRR - 2.0 * fR) + fC;
        Mf = a; if (fabs(b) < fabs(a)) Mf = b;
        dp = (fR - fC) - 0.5 * Mf;
        a = (fR - 2.0 * fC) + fL;
        b = (fC - 2.0 * fL) + fLL;
        Mf = a; if (fabs(b) < fabs(a)) Mf = b;
        dm = (fC - fL) + 0.5 * Mf;
    }
    s = (fabs(f0) > EPS_SIMS) ? fsign1(f0) : 0.0;
    if (s * (fR - fC) < 0.0 && s * (fC - fL) < -(s * (fR - fC))) return dp;
    if (s * (fC - fL) > 0.0 && s * (fR - fC) > -(s * (fC - fL))) return dm;
    return 0.5 * (dp + dm);
}

/* ENO2Boundary, AdvectionMethods.F90:654-707 */
static double eno2_boundary(double f0, double f1, double f2, int dir)
{
    double as = (double)dir;
    double D1 = as * (f1 - f0);
    double D2 = as * (f2 - f1);
    double D3 = as * (f2 - f0) / 2.0;
    if (fabs(D1 - D2) < EPS_SIMS) return D1;
    if (fabs(D1) < fabs(D2)) return (2.0 * D1 + D2) / 3.0;
    if (fabs(D2) < fabs(D1)) return (D1 + 2.0 * D2) / 3.0;
    if (fabs(D1) == fabs(D2)) return (4.0 * D1 + D2) / 5.0;
    if (fabs(D1) < fabs(D3) && fabs(D2) < fabs(D3)) return ((3.0 * D1 + 6.0 * D2) - D3) / 7.0;
    if (fabs(D1) >= fabs(D3) && fabs(D2) >= fabs(D3)) return D3;
    return D1;
}

/* BiasedDerivative, AdvectionMethods.F90:248-302 */
double oracle_biased(double f0, double f1, double f2, double f3, int order, int dir, int enobc)
{
    double as = (double)dir, r = 0.0;
    if (order == 1) r = as * (-(1.0 * f0) + 1.0 * f1);
    if (order == 2) {
        r = as * ((-(3.0 * f0) + 4.0 * f1) - 1.0 * f2) / 2.0;
        if (enobc) r = eno2_boundary(f0, f1, f2, dir);
    }
    if (order == 3) r = as * (((-(11.0 * f0) + 18.0 * f1) - 9.0 * f2) + 2.0 * f3) / 6.0;
    return r;
}

/* GetSignPhi0, AdvectionMethods.F90:576-587 */
static double sign_phi0(double p, double eps)
{
    if (fabs(p) > eps) return fsign1(p);
    return p / eps - (1.0 / PI_REF) * sin((PI_REF * p) / eps);
}

/* deltaf, global_lsm.F90:120-133 */
static double deltaf(double v, double eps)
{
    if (fabs(v) > eps) return 0.0;
    return (1.0 / (2.0 * eps)) * (1.0 + cos((PI_REF * v) / eps));
}

/* GetLambdaMassCorrection, global_lsm.F90:200-219 */
static double lambda_mass(double p, double s, double g, double eps)
{
    double num = -((deltaf(p, eps) * s) * (1.0 - g));
    double del = deltaf(p, eps);
    double den = (del * del) * g;
    if (fabs(den) < EPS_SIMS) return 0.0;
    return num / den;
}

/* ------------------------------------------------------------------------- */
/* R2: rhs_contra_j, rhs_contra_j.F90:17-48                                   */
/* ------------------------------------------------------------------------- */
void oracle_contra(const OBlock *b, const double *u, const double *v, const double *w,
                   double *U1, double *U2, double *U3)
{
    for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
    for (int i = 0; i < b->ni; ++i) {
        size_t l = IX(b, i, j, k);
        double tmp = 1.0 / b->aj[l];
        U1[l] = tmp * ((b->csi[0][l] * u[l] + b->csi[1][l] * v[l]) + b->csi[2][l] * w[l]);
        U2[l] = tmp * ((b->eta[0][l] * u[l] + b->eta[1][l] * v[l]) + b->eta[2][l] * w[l]);
        U3[l] = tmp * ((b->zet[0][l] * u[l] + b->zet[1][l] * v[l]) + b->zet[2][l] * w[l]);
    }
}

/* ------------------------------------------------------------------------- */
/* R4: calc_RH_AD_WENO3, calc_RH_AD_WENO3.F90:86-403 (blanking: 152-188)    */
/* writes interior nodes only; the caller zeroes rh (ad_RKTVD3.F90:55)        */
/* ------------------------------------------------------------------------- */
void oracle_weno3_rhs(const OBlock *b, const double *phi, const double *U1,
                      const double *U2, const double *U3, const int32_t *mask, double *rh)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const size_t si = 1, sj = (size_t)ni, sk = (size_t)ni * nj;
    for (int i = 1; i <= ni - 2; ++i)
    for (int j = 1; j <= nj - 2; ++j)
    for (int k = 1; k <= nk - 2; ++k) {
        size_t l = IX(b, i, j, k);
        if (mask[l] == 0) continue;
        int mb = (i == 1), mf = (i == ni - 2);
        int ml = (j == 1), mr = (j == nj - 2);
        const int md = (k == 1), mu = (k == nk - 2);
        /* blanking: nodes inside a box are skipped, the stencils next to a
           box's i and j faces biased away from it (calc_RH_AD_WENO3.F90:152-188) */
        int inside = 0;
        for (int q = 0; q < b->nblk; ++q) {
            const int32_t *B = b->blank[q];
            if (i >= B[0] && i <= B[1] && j >= B[2] && j <= B[3] && k >= B[4] && k <= B[5]) inside = 1;
            if (i == B[0] - 1 && j >= B[2] && j <= B[3]) mf = 1;
            if (i == B[1] + 1 && j >= B[2] && j <= B[3]) mb = 1;
            if (j == B[2] - 1 && i >= B[0] && i <= B[1]) mr = 1;
            if (j == B[3] + 1 && i >= B[0] && i <= B[1]) ml = 1;
        }
        if (inside) continue;
        const int bnd[3][2] = {{mb, mf}, {ml, mr}, {md, mu}};
        const size_t st[3] = {si, sj, sk};
        const double *Uc[3] = {U1, U2, U3};
        const double dd[3] = {b->dc, b->de, b->dz};
        double L[3];
        for (int d = 0; d < 3; ++d) {
            const size_t s = st[d];
            double ws = Uc[d][l];
            double p0 = phi[l - s], p1 = phi[l], p2 = phi[l + s], p3;
            if (bnd[d][1]) { ws = 1.0; p3 = 0.0; } else p3 = phi[l + 2 * s];
            double fp = oracle_weno3(p0, p1, p2, p3, ws);
            ws = Uc[d][l];
            double q0;
            if (bnd[d][0]) { ws = -1.0; q0 = 0.0; } else q0 = phi[l - 2 * s];
            double fm = oracle_weno3(q0, phi[l - s], phi[l], phi[l + s], ws);
            ws = Uc[d][l];
            L[d] = ws * (dd[d] * (fp - fm));
        }
        rh[l] = (L[0] + L[1]) + L[2];
    }
}

/* ------------------------------------------------------------------------- */
/* R6: bcond_lsm, bcond_lsm.F90:26-287 (faces in order i1, im, j1, jm, k1, km) */
/* ------------------------------------------------------------------------- */
static void bc_face_pair(double *phi, size_t lb, size_t l1, size_t l2, int bt, double big, int is_im)
{
    const double a2 = 4.0 / 3.0, b2 = -1.0 / 3.0, a1 = 1.0, b1 = 0.0;
    if (bt == 0) {
        phi[lb] = phi[l1];
    } else if (bt == 1) {
        double ex = (1.0 + fsign1(fabs(phi[l2]) - big)) / 2.0;
        double ta = a2 + ex * (a1 - a2);
        double tb = b2 + ex * (b1 - b2);
        phi[lb] = ta * phi[l1] + tb * phi[l2];
    } else if (bt == 2 && is_im) {
        phi[lb] = 2.0 * phi[l1] - 1.0 * phi[l2];
    }
}

/* the blanking regions of bcond_lsm.F90:290-562, box by box: the nodes two
   and more inside the box zeroed, then the box's i faces, j faces and the four
   i-j corners extrapolated from outside (second order, first order where the
   farther node is at least BigPhi), each written twice (the face node and
   the node behind it), where the box's blktype is 0 and the face lies at
   least two nodes inside the block.  k faces are not treated (commented out
   in the reference). */
static void blank_extrap(double *phi, size_t lb, size_t lbehind, size_t l1, size_t l2, double big)
{
    const double a2 = 4.0 / 3.0, b2 = -1.0 / 3.0, a1 = 1.0, b1 = 0.0;
    const double ex = (1.0 + fsign1(fabs(phi[l2]) - big)) / 2.0;
    const double ta = a2 + ex * (a1 - a2);
    const double tb = b2 + ex * (b1 - b2);
    phi[lb] = ta * phi[l1] + tb * phi[l2];
    phi[lbehind] = phi[lb];
}

static inline int imax(int a, int c) { return a > c ? a : c; }
static inline int imin(int a, int c) { return a < c ? a : c; }

void oracle_bcond_blanking(const OBlock *b, double *phi)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const double big = b->big_phi;
    for (int q = 0; q < b->nblk; ++q) {
        const int32_t *B = b->blank[q];
        const int ia = B[0], ib = B[1], ja = B[2], jb = B[3], ka = B[4], kb = B[5];
        const int *bt = &B[6];
        for (int k = imax(0, ka + 2); k <= imin(nk - 1, kb - 2); ++k)
        for (int j = imax(0, ja + 2); j <= imin(nj - 1, jb - 2); ++j)
        for (int i = imax(0, ia + 2); i <= imin(ni - 1, ib - 2); ++i) phi[IX(b, i, j, k)] = 0.0;
        const int k0 = imax(0, ka), k1 = imin(nk - 1, kb);
        if (bt[0] == 0 && ia > 1)
            for (int k = k0; k <= k1; ++k) for (int j = imax(0, ja); j <= imin(nj - 1, jb); ++j)
                blank_extrap(phi, IX(b, ia, j, k), IX(b, ia + 1, j, k), IX(b, ia - 1, j, k), IX(b, ia - 2, j, k), big);
        if (bt[1] == 0 && ib < ni - 2)
            for (int k = k0; k <= k1; ++k) for (int j = imax(0, ja); j <= imin(nj - 1, jb); ++j)
                blank_extrap(phi, IX(b, ib, j, k), IX(b, ib - 1, j, k), IX(b, ib + 1, j, k), IX(b, ib + 2, j, k), big);
        if (bt[2] == 0 && ja > 1)
            for (int k = k0; k <= k1; ++k) for (int i = imax(0, ia); i <= imin(ni - 1, ib); ++i)
                blank_extrap(phi, IX(b, i, ja, k), IX(b, i, ja + 1, k), IX(b, i, ja - 1, k), IX(b, i, ja - 2, k), big);
        if (bt[3] == 0 && jb < nj - 2)
            for (int k = k0; k <= k1; ++k) for (int i = imax(0, ia); i <= imin(ni - 1, ib); ++i)
                blank_extrap(phi, IX(b, i, jb, k), IX(b, i, jb - 1, k), IX(b, i, jb + 1, k), IX(b, i, jb + 2, k), big);
        if (bt[0] == 0 && ia > 1 && bt[2] == 0 && ja > 1)
            for (int k = k0; k <= k1; ++k)
                blank_extrap(phi, IX(b, ia, ja, k), IX(b, ia + 1, ja + 1, k), IX(b, ia - 1, ja - 1, k),
                             IX(b, ia - 2, ja - 2, k), big);
        if (bt[1] == 0 && ib < ni - 2 && bt[2] == 0 && ja > 1)
            for (int k = k0; k <= k1; ++k)
                blank_extrap(phi, IX(b, ib, ja, k), IX(b, ib - 1, ja + 1, k), IX(b, ib + 1, ja - 1, k),
                             IX(b, ib + 2, ja - 2, k), big);
        if (bt[1] == 0 && ib < ni - 2 && bt[3] == 0 && jb < nj - 2)
            for (int k = k0; k <= k1; ++k)
                blank_extrap(phi, IX(b, ib, jb, k), IX(b, ib - 1, jb - 1, k), IX(b, ib + 1, jb + 1, k),
                             IX(b, ib + 2, jb + 2, k), big);
        if (bt[0] == 0 && ia > 1 && bt[3] == 0 && jb < nj - 2)
            for (int k = k0; k <= k1; ++k)
                blank_extrap(phi, IX(b, ia, jb, k), IX(b, ia + 1, jb - 1, k), IX(b, ia - 1, jb + 1, k),
                             IX(b, ia - 2, jb + 2, k), big);
    }
}

void oracle_bcond(const OBlock *b, double *phi)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    for (int k = 0; k < nk; ++k) for (int j = 0; j < nj; ++j)
        bc_face_pair(phi, IX(b, 0, j, k), IX(b, 1, j, k), IX(b, 2, j, k), b->btype[0], b->big_phi, 0);
    for (int k = 0; k < nk; ++k) for (int j = 0; j < nj; ++j)
        bc_face_pair(phi, IX(b, ni - 1, j, k), IX(b, ni - 2, j, k), IX(b, ni - 3, j, k), b->btype[1], b->big_phi, 1);
    for (int k = 0; k < nk; ++k) for (int i = 0; i < ni; ++i)
        bc_face_pair(phi, IX(b, i, 0, k), IX(b, i, 1, k), IX(b, i, 2, k), b->btype[2], b->big_phi, 0);
    for (int k = 0; k < nk; ++k) for (int i = 0; i < ni; ++i)
        bc_face_pair(phi, IX(b, i, nj - 1, k), IX(b, i, nj - 2, k), IX(b, i, nj - 3, k), b->btype[3], b->big_phi, 0);
    for (int j = 0; j < nj; ++j) for (int i = 0; i < ni; ++i)
        bc_face_pair(phi, IX(b, i, j, 0), IX(b, i, j, 1), IX(b, i, j, 2), b->btype[4], b->big_phi, 0);
    for (int j = 0; j < nj; ++j) for (int i = 0; i < ni; ++i)
        bc_face_pair(phi, IX(b, i, j, nk - 1), IX(b, i, j, nk - 2), IX(b, i, j, nk - 3), b->btype[5], b->big_phi, 0);
    oracle_bcond_blanking(b, phi);
}

/* ------------------------------------------------------------------------- */
/* calc_RH_AD (orderAD = 0), calc_RH_AD.F90:27-270: first-order upwind        */
/* differences with the velocity split u+ = u/2 + |u/2|, u- = u/2 - |u/2|;    */
/* one-sided (forward at the low face, backward at the high face) times U on  */
/* the faces of every direction.  Per direction the term is assigned (xi) or  */
/* added in front (eta, zeta); nodes with AdvectionNodes <= 0 are left as     */
/* they are.                                                                 */
/* ------------------------------------------------------------------------- */
void oracle_upwind_rhs(const OBlock *b, const double *phi, const double *U1, const double *U2,
                       const double *U3, const int32_t *mask, double *rh)
{
    const int n[3] = {b->ni, b->nj, b->nk};
    const size_t st[3] = {1, (size_t)b->ni, (size_t)b->ni * b->nj};
    const double *Uc[3] = {U1, U2, U3};
    const double dd[3] = {b->dc, b->de, b->dz};
    for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
    for (int i = 0; i < b->ni; ++i) {
        const size_t l = IX(b, i, j, k);
        if (!(mask[l] > 0)) continue;
        const int idx[3] = {i, j, k};
        double r = 0.0;
        for (int d = 0; d < 3; ++d) {
            const size_t s = st[d];
            double t;
            if (idx[d] == 0) {
                t = (dd[d] * Uc[d][l]) * (phi[l + s] - phi[l]);
            } else if (idx[d] == n[d] - 1) {
                t = (dd[d] * Uc[d][l]) * (phi[l] - phi[l - s]);
            } else {
                const double u = 0.5 * Uc[d][l];
                const double up = u + fabs(u), um = u - fabs(u);
                t = dd[d] * (up * (phi[l] - phi[l - s]) + um * (phi[l + s] - phi[l]));
            }
            r = (d == 0) ? t : t + r;
        }
        rh[l] = r;
    }
}

/* ------------------------------------------------------------------------- */
/* R3: ad_RKTVD3, ad_RKTVD3.F90:35-86; order_ad selects calc_RH_AD_WENO3 (1)  */
/* or calc_RH_AD (0), ad_RKTVD3.F90:57-58                                     */
/* ------------------------------------------------------------------------- */
void oracle_rk3_order(const OBlock *b, double *phi_np1, const double *phi_n, const int32_t *mask,
                      const double *U1, const double *U2, const double *U3, double dt, int order_ad)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    const double c[3][3] = {{1.0, 0.0, 1.0},
                            {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                            {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    double *rh = (double *)malloc(n * sizeof(double));
    memcpy(phi_np1, phi_n, n * sizeof(double));
    for (int s = 0; s < 3; ++s) {
        memset(rh, 0, n * sizeof(double));
        if (order_ad == 0) oracle_upwind_rhs(b, phi_np1, U1, U2, U3, mask, rh);
        else oracle_weno3_rhs(b, phi_np1, U1, U2, U3, mask, rh);
        for (size_t l = 0; l < n; ++l)
            if (mask[l] > 0)
                phi_np1[l] = (c[s][0] * phi_n[l] + c[s][1] * phi_np1[l])
                           - ((c[s][2] * b->aj[l]) * dt) * rh[l];
        oracle_bcond(b, phi_np1);
    }
    free(rh);
}

void oracle_rk3(const OBlock *b, double *phi_np1, const double *phi_n, const int32_t *mask,
                const double *U1, const double *U2, const double *U3, double dt)
{
    oracle_rk3_order(b, phi_np1, phi_n, mask, U1, U2, U3, dt, 1);
}

/* ------------------------------------------------------------------------- */
/* R10: calc_NormGrad_d, calc_NormGrad_d.F90:39-2446                          */
/* ------------------------------------------------------------------------- */
void oracle_normgrad(const OBlock *b, const double *d, const double *phi0, double *ng,
                     int order_bc, int eno_bc)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const int n[3] = {ni, nj, nk};
    const size_t st[3] = {1, (size_t)ni, (size_t)ni * nj};
    const double dd[3] = {b->dc, b->de, b->dz};
    /* i = 1 / i = im face blocks never assign phi0, so they use the value left
       by the last interior node of the interior loop (calc_NormGrad_d.F90:91,
       187-367): node (im-1, jm-1, km-1). */
    const double phi0_stale = phi0[IX(b, ni - 2, nj - 2, nk - 2)];
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const int idx[3] = {i, j, k};
        const size_t l = IX(b, i, j, k);
        int on[3];
        for (int a = 0; a < 3; ++a) on[a] = (idx[a] == 0 || idx[a] == n[a] - 1);
        const double p0 = (on[0] && !on[1] && !on[2]) ? phi0_stale : phi0[l];
        double g[3];
        for (int a = 0; a < 3; ++a) {
            const size_t s = st[a];
            if (on[a]) {
                const int dir = (idx[a] == 0) ? 1 : -1;
                if (a == 0 && i == 0 && k == 0 && !on[1]) {
                    /* edge i = 1, k = 1 (calc_NormGrad_d.F90:1281-1283) */
                    const double dc2 = 0.5 * b->dc;
                    g[a] = dc2 * ((-(1.0 * d[l + 2]) + 4.0 * d[l + 1]) - 3.0 * d[l]);
                } else {
                    const long ss = (long)s * dir;
                    g[a] = dd[a] * oracle_biased(d[l], d[l + ss], d[l + 2 * ss], d[l + 3 * ss],
                                                 order_bc, dir, eno_bc);
                }
            } else {
                const int lo = (idx[a] == 1), hi = (idx[a] == n[a] - 2);
                const double fLL = lo ? 0.0 : d[l - 2 * s];
                const double fRR = hi ? 0.0 : d[l + 2 * s];
                g[a] = dd[a] * oracle_eno2(p0, fLL, d[l - s], d[l], d[l + s], fRR, lo, hi);
            }
        }
        const double gx = (b->csi[0][l] * g[0] + b->eta[0][l] * g[1]) + b->zet[0][l] * g[2];
        const double gy = (b->csi[1][l] * g[0] + b->eta[1][l] * g[1]) + b->zet[1][l] * g[2];
        const double gz = (b->csi[2][l] * g[0] + b->eta[2][l] * g[1]) + b->zet[2][l] * g[2];
        ng[l] = sqrt((gx * gx + gy * gy) + gz * gz);
    }
}

/* ------------------------------------------------------------------------- */
/* R8: reinitialisation_test2, reinitialisation_test2.F90:42-210 (non-hybrid) */
/* returns the number of pseudo-time iterations that ran                      */
/* ------------------------------------------------------------------------- */
int oracle_reinit_eno2(const OBlock *b, const OReinit *p, const double *phi0, double *out)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    const double c[3][3] = {{1.0, 0.0, 1.0},
                            {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                            {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    double *s = (double *)malloc(n * sizeof(double));
    double *dn = (double *)malloc(n * sizeof(double));
    double *ng = (double *)malloc(n * sizeof(double));
    double *ngc = (double *)malloc(n * sizeof(double));
    for (size_t l = 0; l < n; ++l) s[l] = sign_phi0(phi0[l], p->epslsm);
    memcpy(out, phi0, n * sizeof(double));
    oracle_normgrad(b, out, phi0, ng, p->order_bc, p->eno_bc);
    int it;
    for (it = 0; it < p->rn_itermax; ++it) {
        double mx = ng[0], mn = ng[0];
        for (size_t l = 1; l < n; ++l) { if (ng[l] > mx) mx = ng[l]; if (ng[l] < mn) mn = ng[l]; }
        if (fabs(mx - 1.0) < p->eps_reinit && fabs(mn - 1.0) < p->eps_reinit) break;
        memcpy(dn, out, n * sizeof(double));
        oracle_normgrad(b, dn, phi0, ngc, p->order_bc, p->eno_bc);
        for (int r = 0; r < 3; ++r) {
            oracle_normgrad(b, out, phi0, ng, p->order_bc, p->eno_bc);
            for (size_t l = 0; l < n; ++l) {
                if (ngc[l] - 1.0 > p->eps_reinit) {
                    double rhs = s[l] * (ng[l] - 1.0);
                    double lam = lambda_mass(out[l], s[l], ng[l], p->epslsm);
                    double aux = deltaf(out[l], p->epslsm) * ng[l];
                    rhs = rhs - lam * aux;
                    out[l] = (c[r][0] * dn[l] + c[r][1] * out[l]) - (c[r][2] * p->rn_timestep) * rhs;
                }
            }
        }
    }
    free(s); free(dn); free(ng); free(ngc);
    return it;
}

/* ------------------------------------------------------------------------- */
/* one-sided differences with the boundary copies of calc_sign_ini.F90:41-176 */
/* and calc_RH_RN.F90:35-175; Godunov |grad| of calc_sign_ini.F90:199-250     */
/* ------------------------------------------------------------------------- */
static void godunov_norm(const OBlock *b, const double *f, size_t l, int i, int j, int k,
                         double *ng_plus, double *ng_minus)
{
    const int n[3] = {b->ni, b->nj, b->nk};
    const int idx[3] = {i, j, k};
    const size_t st[3] = {1, (size_t)b->ni, (size_t)b->ni * b->nj};
    const double dd[3] = {b->dc, b->de, b->dz};
    double mas[3], men[3];
    for (int a = 0; a < 3; ++a) {
        const size_t s = st[a];
        if (idx[a] == 0) {
            mas[a] = dd[a] * (f[l + s] - f[l]);
            men[a] = mas[a];
        } else if (idx[a] == n[a] - 1) {
            men[a] = dd[a] * (f[l] - f[l - s]);
            mas[a] = men[a];
        } else {
            mas[a] = dd[a] * (f[l + s] - f[l]);
            men[a] = dd[a] * (f[l] - f[l - s]);
        }
    }
    double db[3], df[3];
    for (int c = 0; c < 3; ++c) {
        db[c] = (b->csi[c][l] * men[0] + b->eta[c][l] * men[1]) + b->zet[c][l] * men[2];
        df[c] = (b->csi[c][l] * mas[0] + b->eta[c][l] * mas[1]) + b->zet[c][l] * mas[2];
    }
    double bm[3], bn[3], fm[3], fn[3];
    for (int c = 0; c < 3; ++c) {
        double a = db[c] / 2.0; bm[c] = a + fabs(a); bn[c] = a - fabs(a);
        a = df[c] / 2.0;        fm[c] = a + fabs(a); fn[c] = a - fabs(a);
    }
#define MAX2(p, q) ((p) > (q) ? (p) : (q))
    *ng_plus = sqrt((MAX2(bm[0] * bm[0], fn[0] * fn[0]) + MAX2(bm[1] * bm[1], fn[1] * fn[1]))
                    + MAX2(bm[2] * bm[2], fn[2] * fn[2]));
    *ng_minus = sqrt((MAX2(bn[0] * bn[0], fm[0] * fm[0]) + MAX2(bn[1] * bn[1], fm[1] * fm[1]))
                     + MAX2(bn[2] * bn[2], fm[2] * fm[2]));
#undef MAX2
}

/* R13: calc_sign_ini, calc_sign_ini.F90:1-280 */
void oracle_sign_ini(const OBlock *b, const double *phi0, double *sgn, double *ng,
                     double epslsm, int m_sphi0)
{
    for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
    for (int i = 0; i < b->ni; ++i) {
        size_t l = IX(b, i, j, k);
        double gp, gm, g;
        godunov_norm(b, phi0, l, i, j, k, &gp, &gm);
        const double p = phi0[l];
        if (p > 0.0) g = gp; else if (p < 0.0) g = gm; else g = 1.0;
        if (m_sphi0 == 1) {
            if (p >= epslsm) sgn[l] = 1.0;
            else if (p <= (-1.0) * epslsm) sgn[l] = -1.0;
            else sgn[l] = p / epslsm - (1.0 / PI_REF) * sin((PI_REF * p) / epslsm);
        } else {
            sgn[l] = p / sqrt(p * p + (g * epslsm) * (g * epslsm));
        }
        ng[l] = g;
    }
}

/* R14: calc_RH_RN, calc_RH_RN.F90:1-240 */
void oracle_rh_rn(const OBlock *b, const double *d, const double *sgn, double *rh)
{
    for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
    for (int i = 0; i < b->ni; ++i) {
        size_t l = IX(b, i, j, k);
        double gp, gm, g;
        godunov_norm(b, d, l, i, j, k, &gp, &gm);
        if (sgn[l] > 0.0) g = gp; else if (sgn[l] < 0.0) g = gm; else g = 1.0;
        rh[l] = sgn[l] * (g - 1.0);
    }
}

/* R15: calc_lambda_RN, method 0, calc_lambda_RN.F90:42-63 */
void oracle_lambda_rn(const OBlock *b, const double *phi0, const double *ng0, const double *d,
                      double *lam, double dt, double epslsm)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    for (size_t l = 0; l < n; ++l) {
        double vol = ((1.0 / b->dc) * (1.0 / b->de)) * (1.0 / b->dz);
        double Lop = deltaf(phi0[l], epslsm) * ((1.0 / dt) * (d[l] - phi0[l]));
        double vn = vol * Lop;
        double del = deltaf(phi0[l], epslsm);
        double vd = (vol * (del * del)) * ng0[l];
        lam[l] = (vd == 0.0) ? 0.0 : -(vn / vd);
    }
}

/* testfilter, testfilter.F90:1-124: at the nodes off the block's faces the
   average of the 27-node neighbourhood with weights 64 (the node), 16 (the six
   face neighbours), 4 (the twelve edge neighbours) and 1 (the eight corners),
   each divided by aj there, summed in the reference's order; the face nodes
   keep var */
static void testfilter(const OBlock *b, const double *var, double *varf)
{
    static const int off[27][3] = {
        {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
        {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
        {-1, 0, -1}, {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},   {1, 1, -1},
        {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    memcpy(varf, var, n * sizeof(double));
    for (int k = 1; k < b->nk - 1; ++k)
    for (int j = 1; j < b->nj - 1; ++j)
    for (int i = 1; i < b->ni - 1; ++i) {
        double avgvar = 0.0, avgvol = 0.0;
        for (int m = 0; m < 27; ++m) {
            const int nz = (off[m][0] != 0) + (off[m][1] != 0) + (off[m][2] != 0);
            const double c = nz == 0 ? 64.0 : nz == 1 ? 16.0 : nz == 2 ? 4.0 : 1.0;
            const size_t q = IX(b, i + off[m][0], j + off[m][1], k + off[m][2]);
            const double w = c / b->aj[q];
            avgvar = avgvar + w * var[q];
            avgvol = avgvol + w;
        }
        varf[IX(b, i, j, k)] = avgvar / avgvol;
    }
}

/* R15: calc_lambda_RN, method 1 (the filtered quotient), calc_lambda_RN.F90:65-120:
   numerator and denominator filtered by testfilter, lambda = 0 where the
   filtered denominator is within 1E-10 (a default-real literal) of 0 */
void oracle_lambda_rn_simpson(const OBlock *b, const double *phi0, const double *ng0, const double *d,
                              double *lam, double dt, double epslsm)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    if (n == 0) return;
    double *num = (double *)malloc(n * sizeof(double)), *den = (double *)malloc(n * sizeof(double));
    double *fnum = (double *)malloc(n * sizeof(double)), *fden = (double *)malloc(n * sizeof(double));
    for (size_t l = 0; l < n; ++l) {
        const double del = deltaf(phi0[l], epslsm);
        num[l] = del * ((1.0 / dt) * (d[l] - phi0[l]));
        den[l] = (del * del) * ng0[l];
    }
    testfilter(b, num, fnum);
    testfilter(b, den, fden);
    const double tiny = (double)1e-10f;
    for (size_t l = 0; l < n; ++l) lam[l] = (fabs(fden[l]) <= tiny) ? 0.0 : -(fnum[l] / fden[l]);
    free(num); free(den); free(fnum); free(fden);
}

/* R12: reinitialization_benchmark, reinitialization_benchmark.F90:45-222    */
/* (M_Grad = 0, no obstacles, non-hybrid)                                    */
void oracle_reinit_godunov(const OBlock *b, const OReinit *p, int m_sphi0, int m_ti,
                           const double *phi0, double *out)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    const double c[3][3] = {{1.0, 0.0, 1.0},
                            {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                            {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    const double dt = p->rn_timestep;
    double *sn = (double *)malloc(n * sizeof(double));
    double *sg = (double *)malloc(n * sizeof(double));
    double *ng0 = (double *)malloc(n * sizeof(double));
    double *rh = (double *)malloc(n * sizeof(double));
    double *lam = (double *)malloc(n * sizeof(double));
    memcpy(sn, phi0, n * sizeof(double));
    memcpy(out, phi0, n * sizeof(double));
    oracle_sign_ini(b, phi0, sg, ng0, p->epslsm, m_sphi0);
    for (int it = 0; it < p->rn_itermax; ++it) {
        if (m_ti == 1) {
            for (int r = 0; r < 3; ++r) {
                oracle_rh_rn(b, out, sg, rh);
                for (size_t l = 0; l < n; ++l)
                    out[l] = (c[r][0] * sn[l] + c[r][1] * out[l]) - (c[r][2] * dt) * rh[l];
            }
        } else {
            oracle_rh_rn(b, out, sg, rh);
            for (size_t l = 0; l < n; ++l) out[l] = sn[l] - dt * rh[l];
        }
        if (p->lambda_method == 1) oracle_lambda_rn_simpson(b, phi0, ng0, out, lam, dt, p->epslsm);
        else oracle_lambda_rn(b, phi0, ng0, out, lam, dt, p->epslsm);
        for (size_t l = 0; l < n; ++l)
            out[l] = out[l] + ((dt * lam[l]) * deltaf(phi0[l], p->epslsm)) * ng0[l];
        memcpy(sn, out, n * sizeof(double));
    }
    free(sn); free(sg); free(ng0); free(rh); free(lam);
}

/* ------------------------------------------------------------------------- */
/* R16: TotalWaterVolume + TetrahedronMethods                                */
/* ------------------------------------------------------------------------- */

/* M44Det, TetrahedronMethods.F90:1698-1720 (A is row-major 4x4, A[r][c]) */
static double m44det(const double A[4][4])
{
#define a(r, c) A[(r) - 1][(c) - 1]
    return ((a(1,1)*((a(2,2)*(a(3,3)*a(4,4)-a(3,4)*a(4,3))+a(2,3)*(a(3,4)*a(4,2)-a(3,2)*a(4,4)))+a(2,4)*(a(3,2)*a(4,3)-
            a(3,3)*a(4,2)))-a(1,2)*((a(2,1)*(a(3,3)*a(4,4)-a(3,4)*a(4,3))+a(2,3)*(a(3,4)*a(4,1)-a(3,1)*a(4,4)))+
            a(2,4)*(a(3,1)*a(4,3)-a(3,3)*a(4,1))))+a(1,3)*((a(2,1)*(a(3,2)*a(4,4)-a(3,4)*a(4,2))+a(2,2)*(a(3,4)*a(4,1)-
            a(3,1)*a(4,4)))+a(2,4)*(a(3,1)*a(4,2)-a(3,2)*a(4,1))))-a(1,4)*((a(2,1)*(a(3,2)*a(4,3)-a(3,3)*a(4,2))+
            a(2,2)*(a(3,3)*a(4,1)-a(3,1)*a(4,3)))+a(2,3)*(a(3,1)*a(4,2)-a(3,2)*a(4,1)));
#undef a
}

/* GetTetrahedronVolume, TetrahedronMethods.F90:666-688 */
static double tet_volume(const double v[4][3])
{
    double A[4][4];
    for (int r = 0; r < 4; ++r) { A[r][0] = v[r][0]; A[r][1] = v[r][1]; A[r][2] = v[r][2]; A[r][3] = 1.0; }
    return (1.0 / 6.0) * fabs(m44det(A));
}

static void cross3(const double *u, const double *v, double *o)
{
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
}

static double dot3(const double *u, const double *v) { return (u[0] * v[0] + u[1] * v[1]) + u[2] * v[2]; }

/* HexahedronVolume (Davies & Salmond), TetrahedronMethods.F90:1724-1793 */
static double hex_volume(const double V[8][3])
{
    const double *A = V[0], *B = V[1], *C = V[2], *D = V[3], *E = V[4], *F = V[5], *G = V[6], *H = V[7];
    double AC[3], AF[3], AH[3], BE[3], CH[3], DB[3], ED[3], FC[3];
    double GA[3], GB[3], GC[3], GD[3], GE[3], GF[3], GH[3], HF[3];
    for (int c = 0; c < 3; ++c) {
        AC[c] = C[c] - A[c]; AF[c] = F[c] - A[c]; AH[c] = H[c] - A[c];
        BE[c] = E[c] - B[c]; CH[c] = H[c] - C[c]; DB[c] = B[c] - D[c];
        ED[c] = D[c] - E[c]; FC[c] = C[c] - F[c];
        GA[c] = A[c] - G[c]; GB[c] = B[c] - G[c]; GC[c] = C[c] - G[c]; GD[c] = D[c] - G[c];
        GE[c] = E[c] - G[c]; GF[c] = F[c] - G[c]; GH[c] = H[c] - G[c]; HF[c] = F[c] - H[c];
    }
    double t[3], s[3], acc;
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GB[c];
    cross3(DB, AC, t); acc = dot3(s, t);
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GE[c];
    cross3(BE, AF, t); acc = acc + dot3(s, t);
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GD[c];
    cross3(ED, AH, t); acc = acc + dot3(s, t);
    cross3(HF, GE, t); acc = acc + dot3(GF, t);
    cross3(CH, GD, t); acc = acc + dot3(GH, t);
    cross3(FC, GB, t); acc = acc + dot3(GC, t);
    return -((1.0 / 12.0) * acc);
}

/* edges of a tetrahedron (0-based vertices), TetrahedronMethods.F90:71-79 */
static const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
/* edge-intersection flags per case, TetrahedronMethods.F90:130-135 */
static const int TET_EDGE_FLAGS[16] = {0x00, 0x0d, 0x13, 0x1e, 0x26, 0x2b, 0x35, 0x38,
                                      0x38, 0x35, 0x2b, 0x26, 0x1e, 0x13, 0x0d, 0x00};

/* GetWaterVolume, TetrahedronMethods.F90:710-954 */
static double water_volume_cut(const double v[4][3], const double E[6][3], int flag)
{
    const int cs = flag + 1;
    const double T = tet_volume(v);
    double t1[4][3], t2[4][3], t3[4][3];
#define CP(dst, src) do { (dst)[0] = (src)[0]; (dst)[1] = (src)[1]; (dst)[2] = (src)[2]; } while (0)
    switch (cs) {
    case 1: return T;
    case 16: return 0.0;
    case 2: case 15:
        CP(t1[0], v[0]); CP(t1[1], E[0]); CP(t1[2], E[2]); CP(t1[3], E[3]);
        return cs == 2 ? T - tet_volume(t1) : tet_volume(t1);
    case 3: case 14:
        CP(t1[0], v[1]); CP(t1[1], E[0]); CP(t1[2], E[1]); CP(t1[3], E[4]);
        return cs == 3 ? T - tet_volume(t1) : tet_volume(t1);
    case 5: case 12:
        CP(t1[0], v[2]); CP(t1[1], E[1]); CP(t1[2], E[2]); CP(t1[3], E[5]);
        return cs == 5 ? T - tet_volume(t1) : tet_volume(t1);
    case 8: case 9:
        CP(t1[0], v[3]); CP(t1[1], E[3]); CP(t1[2], E[4]); CP(t1[3], E[5]);
        return cs == 8 ? tet_volume(t1) : T - tet_volume(t1);
    case 4: case 13: {
        CP(t1[0], v[0]); CP(t1[1], E[2]); CP(t1[2], E[3]); CP(t1[3], E[4]);
        CP(t2[0], v[0]); CP(t2[1], E[1]); CP(t2[2], E[2]); CP(t2[3], E[4]);
        CP(t3[0], v[0]); CP(t3[1], v[1]); CP(t3[2], E[1]); CP(t3[3], E[4]);
        double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 4 ? T - aux : aux;
    }
    case 6: case 11: {
        CP(t1[0], v[0]); CP(t1[1], E[0]); CP(t1[2], E[1]); CP(t1[3], E[5]);
        CP(t2[0], v[0]); CP(t2[1], E[0]); CP(t2[2], E[3]); CP(t2[3], E[5]);
        CP(t3[0], v[0]); CP(t3[1], v[2]); CP(t3[2], E[1]); CP(t3[3], E[5]);
        double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 6 ? T - aux : aux;
    }
    case 7: case 10: {
        CP(t1[0], v[1]); CP(t1[1], E[0]); CP(t1[2], E[2]); CP(t1[3], E[5]);
        CP(t2[0], v[1]); CP(t2[1], E[0]); CP(t2[2], E[4]); CP(t2[3], E[5]);
        CP(t3[0], v[1]); CP(t3[1], v[2]); CP(t3[2], E[2]); CP(t3[3], E[5]);
        double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 7 ? T - aux : aux;
    }
    }
#undef CP
    return 0.0;
}

/* water volume of MarchingTetrahedron, TetrahedronMethods.F90:26-664 */
static double marching_tet_water(const double v[4][3], const double ph[4])
{
    const double tiny = 2.2250738585072014e-308;   /* tiny(eps_sims) */
    int sg[4], npos = 0, nneg = 0, nzero = 0;
    for (int q = 0; q < 4; ++q) {
        sg[q] = 0;
        if (ph[q] > tiny) sg[q] = 1;
        if (ph[q] < -tiny) sg[q] = -1;
        npos += sg[q] == 1; nneg += sg[q] == -1; nzero += sg[q] == 0;
    }
    if (npos == 4) return tet_volume(v);
    if (nneg == 4) return 0.0;
    int flag = 0;
    if (nzero > 0 && (nneg == 0 || npos == 0)) {
        /* isosurface through 1, 2 or 3 vertices: the tetrahedron is full when
           the first listed non-zero vertex of the matching case is positive */
        if (nzero == 1) {
            if (sg[0] == 0) return sg[1] == 1 ? tet_volume(v) : 0.0;
            if (sg[1] == 0) return sg[0] == 1 ? tet_volume(v) : 0.0;
            if (sg[2] == 0) return sg[0] == 1 ? tet_volume(v) : 0.0;
            return sg[0] == 1 ? tet_volume(v) : 0.0;
        }
        if (nzero == 2) {
            if (sg[0] == 0 && sg[1] == 0) return sg[2] == 1 ? tet_volume(v) : 0.0;
            if (sg[1] == 0 && sg[2] == 0) return sg[0] == 1 ? tet_volume(v) : 0.0;
            if (sg[2] == 0 && sg[0] == 0) return sg[1] == 1 ? tet_volume(v) : 0.0;
            if (sg[0] == 0 && sg[3] == 0) return sg[1] == 1 ? tet_volume(v) : 0.0;
            if (sg[1] == 0 && sg[3] == 0) return sg[0] == 1 ? tet_volume(v) : 0.0;
            return sg[0] == 1 ? tet_volume(v) : 0.0;
        }
        if (nzero == 3) {
            for (int q = 0; q < 4; ++q)
                if (sg[q] != 0) return sg[q] == 1 ? tet_volume(v) : 0.0;
        }
        /* four zero vertices: no case matches, FlagIndex stays 0 */
        return tet_volume(v);
    }
    for (int q = 0; q < 4; ++q) if (sg[q] == -1) flag |= 1 << q;
    if (flag == 0) return tet_volume(v);
    if (flag == 15) return 0.0;
    const int ef = TET_EDGE_FLAGS[flag];
    double E[6][3];
    for (int e = 0; e < 6; ++e) { E[e][0] = E[e][1] = E[e][2] = -9.9; }
    for (int e = 0; e < 6; ++e) {
        if (!(ef & (1 << e))) continue;
        const int s0 = TET_EDGE[e][0], s1 = TET_EDGE[e][1];
        const double den = ph[s1] - ph[s0];
        double s = 0.0;
        if (!(fabs(den) < EPS_SIMS * EPS_SIMS * EPS_SIMS)) s = ph[s1] / den;
        for (int c = 0; c < 3; ++c) E[e][c] = v[s1][c] + s * (v[s0][c] - v[s1][c]);
    }
    return water_volume_cut(v, E, flag);
}

/* tetrahedra of a cell, TriangulationCase 1, TetrahedronMethods.F90:1929-1936 */
static const int CELL_TETS[6][4] = {{0, 5, 1, 6}, {0, 1, 2, 6}, {0, 2, 3, 6},
                                    {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}};

/* CellAnalysis, TetrahedronMethods.F90:1823-2121 (volumes only) */
static void cell_analysis(const double ph[8], const double V[8][3], double *pc_vol, double *st_vol)
{
    int pcell = 0;
    *pc_vol = 0.0; *st_vol = 0.0;
    for (int q = 0; q < 8; ++q) if (fabs(ph[q]) < EPS_SIMS) pcell = 1;
    if (!pcell) {
        double sgn = fsign1(ph[0]);
        for (int q = 1; q < 8; ++q) if (sgn * ph[q] < 0.0) { pcell = 1; break; }
    }
    if (!pcell) return;
    for (int t = 0; t < 6; ++t) {
        int pct = 0, second = 1;
        double pa = ph[CELL_TETS[t][0]];
        if (pa <= EPS_SIMS) second = 0;
        if (fabs(pa) < EPS_SIMS) pct = 1;
        double sgn = fsign1(pa);
        for (int q = 1; q < 4; ++q) {
            pa = ph[CELL_TETS[t][q]];
            if (sgn * pa < 0.0 || fabs(pa) < EPS_SIMS) {
                if (pa <= EPS_SIMS) second = 0;
                pct = 1;
            }
        }
        double tv[4][3], tp[4];
        for (int q = 0; q < 4; ++q) {
            const int vq = CELL_TETS[t][q];
            tv[q][0] = V[vq][0]; tv[q][1] = V[vq][1]; tv[q][2] = V[vq][2];
            tp[q] = ph[vq];
        }
        if (pct) *pc_vol = *pc_vol + marching_tet_water(tv, tp);
        if (second && !pct) *st_vol = *st_vol + tet_volume(tv);
    }
}

/* TotalWaterVolume, TotalWaterVolume.F90:69-294 (blanking: 139-207).  vol[0..2] =
   (K-tetrahedra, second tetrahedra, bulk cells); mask = AdvectionNodes. */
void oracle_water_volume(const OBlock *b, const double *phi, double vol[3], int32_t *mask)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const size_t n = (size_t)ni * nj * nk;
    if (mask) memset(mask, 0, n * sizeof(int32_t));
    double vk = 0.0, vs = 0.0, vb = 0.0;
    static const int OFF[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
    for (int k = 0; k <= nk - 2; ++k)
    for (int j = 0; j <= nj - 2; ++j)
    for (int i = 0; i <= ni - 2; ++i) {
        /* blanking: a node inside a box is neither an interface node nor the
           corner of a counted cell (TotalWaterVolume.F90:139-160) */
        int blanked = 0;
        for (int q = 0; q < b->nblk; ++q) {
            const int32_t *B = b->blank[q];
            if (i >= B[0] && i <= B[1] && j >= B[2] && j <= B[3] && k >= B[4] && k <= B[5]) blanked = 1;
        }
        if (blanked) continue;
        const int km = k - 1 < 0 ? 0 : k - 1, kp = k + 1 > nk - 1 ? nk - 1 : k + 1;
        const int jm = j - 1 < 0 ? 0 : j - 1, jp = j + 1 > nj - 1 ? nj - 1 : j + 1;
        const int im = i - 1 < 0 ? 0 : i - 1, ip = i + 1 > ni - 1 ? ni - 1 : i + 1;
        const double p = phi[IX(b, i, j, k)];
        const double sgn = fsign1(p);
        int change = fabs(p) < EPS_SIMS;
        if (!change) {
            for (int kk = km; kk <= kp && !change; ++kk)
            for (int jj = jm; jj <= jp && !change; ++jj)
            for (int ii = im; ii <= ip; ++ii) {
                /* nodes strictly inside a box are not searched (TotalWaterVolume.F90:190-207) */
                int skip = 0;
                for (int q = 0; q < b->nblk; ++q) {
                    const int32_t *B = b->blank[q];
                    if (ii > B[0] && ii < B[1] && jj > B[2] && jj < B[3] && kk > B[4] && kk < B[5]) skip = 1;
                }
                if (skip) continue;
                const double q = phi[IX(b, ii, jj, kk)];
                if (sgn * q < 0.0 || fabs(q) < EPS_SIMS) { change = 1; break; }
            }
        }
        if (change && mask) {
            for (int kk = km; kk <= kp; ++kk)
            for (int jj = jm; jj <= jp; ++jj)
            for (int ii = im; ii <= ip; ++ii) mask[IX(b, ii, jj, kk)] = 1;
        }
        double V[8][3], ph[8];
        int allp = 1, alln = 1;
        for (int q = 0; q < 8; ++q) {
            const size_t l = IX(b, i + OFF[q][0], j + OFF[q][1], k + OFF[q][2]);
            V[q][0] = b->x[l]; V[q][1] = b->y[l]; V[q][2] = b->z[l];
            ph[q] = phi[l];
            if (!(ph[q] > EPS_SIMS)) allp = 0;
            if (!(ph[q] < -EPS_SIMS)) alln = 0;
        }
        double cw = 0.0, st = 0.0, pc = 0.0;
        if (allp) cw = hex_volume(V);
        else if (alln) continue;
        else cell_analysis(ph, V, &pc, &st);
        vk = vk + pc;
        vs = vs + st;
        vb = vb + cw;
    }
    vol[0] = vk; vol[1] = vs; vol[2] = vb;
}

/* ------------------------------------------------------------------------- */
/* Caller-side consumers of phi (mg_driver.F90:106-150).                      */
/* calc_phi_gradient_ENO2.F90:1-2025 (nc = 3: gradient of phi) and            */
/* calc_h_gradient_ENO2.F90:1-2018 (nc = 2: x and y components of grad h):   */
/* the calc_NormGrad_d stencils with the node's own value as the ENO2         */
/* upwind sign (GetENO2Reconstruction(phiC, ...), calc_phi_gradient_ENO2.F90: */
/* 178,195,212) and no stale-phi0 face quirk; interior, face and edge nodes   */
/* are written, the eight vertices are not (the routines have no vertex       */
/* blocks: their outputs are intent(inout) and keep the caller's values).    */
/* The edge i = 1, k = 1 keeps the fixed second-order xi formula              */
/* (calc_phi_gradient_ENO2.F90:1413-1415, calc_h_gradient_ENO2.F90:1406).    */
/* ------------------------------------------------------------------------- */
void oracle_gradient_eno2(const OBlock *b, const double *f, double *g0, double *g1, double *g2,
                          int nc, int order_bc, int eno_bc)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const int n[3] = {ni, nj, nk};
    const size_t st[3] = {1, (size_t)ni, (size_t)ni * nj};
    const double dd[3] = {b->dc, b->de, b->dz};
    double *out[3] = {g0, g1, g2};
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const int idx[3] = {i, j, k};
        const size_t l = IX(b, i, j, k);
        int on[3];
        for (int a = 0; a < 3; ++a) on[a] = (idx[a] == 0 || idx[a] == n[a] - 1);
        if (on[0] && on[1] && on[2]) continue;
        double g[3];
        /* blanking, interior loop only (calc_phi_gradient_ENO2.F90:133-158,
           216-231; calc_h_gradient_ENO2.F90 the same): the stencils one node
           off a box's i or j face are biased away from it, and the xi (eta)
           derivative is zero on the box's i (j) faces */
        int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0}, zero_d[2] = {0, 0};
        if (!on[0] && !on[1] && !on[2])
            for (int q = 0; q < b->nblk; ++q) {
                const int32_t *B = b->blank[q];
                const int jin = j >= B[2] && j <= B[3], iin = i >= B[0] && i <= B[1];
                if (i == B[0] - 1 && jin) bhi[0] = 1;
                if (i == B[1] + 1 && jin) blo[0] = 1;
                if (j == B[2] - 1 && iin) bhi[1] = 1;
                if (j == B[3] + 1 && iin) blo[1] = 1;
                if ((i == B[0] || i == B[1]) && jin) zero_d[0] = 1;
                if ((j == B[2] || j == B[3]) && iin) zero_d[1] = 1;
            }
        for (int a = 0; a < 3; ++a) {
            const size_t s = st[a];
            if (on[a]) {
                const int dir = (idx[a] == 0) ? 1 : -1;
                if (a == 0 && i == 0 && k == 0 && !on[1]) {
                    const double dc2 = 0.5 * b->dc;
                    g[a] = dc2 * ((-(1.0 * f[l + 2]) + 4.0 * f[l + 1]) - 3.0 * f[l]);
                } else {
                    const long ss = (long)s * dir;
                    g[a] = dd[a] * oracle_biased(f[l], f[l + ss], f[l + 2 * ss], f[l + 3 * ss],
                                                 order_bc, dir, eno_bc);
                }
            } else {
                const int lo = (idx[a] == 1) || blo[a], hi = (idx[a] == n[a] - 2) || bhi[a];
                const double fLL = lo ? 0.0 : f[l - 2 * s];
                const double fRR = hi ? 0.0 : f[l + 2 * s];
                g[a] = dd[a] * oracle_eno2(f[l], fLL, f[l - s], f[l], f[l + s], fRR, lo, hi);
            }
        }
        if (zero_d[0]) g[0] = 0.0;
        if (zero_d[1]) g[1] = 0.0;
        for (int c = 0; c < nc; ++c)
            out[c][l] = (b->csi[c][l] * g[0] + b->eta[c][l] * g[1]) + b->zet[c][l] * g[2];
    }
}

/* find_z_free_surface.F90:1-62: the first k (from the bottom) with          */
/* phi(k) * phi(k+1) <= 0 and the free-surface height in that interval.      */
/* Returns 0 when no such interval exists.                                    */
static int find_z_free_surface(const double *z, const double *phi, size_t sk, int nk, double *zfs)
{
    for (int k = 0; k < nk - 1; ++k) {
        const double p0 = phi[k * sk], p1 = phi[(k + 1) * sk];
        if (p0 * p1 <= 0.0) {
            const double z0 = z[k * sk], z1 = z[(k + 1) * sk];
            if (fabs(p0) < EPS_SIMS && fabs(p1) > EPS_SIMS) *zfs = z0;
            else if (fabs(p1) < EPS_SIMS && fabs(p0) > EPS_SIMS) *zfs = z1;
            else if (fabs(p1) < EPS_SIMS && fabs(p0) < EPS_SIMS) *zfs = 0.5 * (z1 + z0);
            else *zfs = z0 + ((z1 - z0) * (-p0)) / (p1 - p0);
            return 1;
        }
    }
    return 0;
}

/* calc_h.F90:101-185: h(i,j,k) = z_fs(i,j) - z(i,j,k) over every column,     */
/* h = 0 in a column without a free surface (the reference prints 'free       */
/* surface not found' there).  Returns the number of such columns.            */
int oracle_calc_h(const OBlock *b, const double *z, const double *phi, double *h)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const size_t sk = (size_t)ni * nj;
    int missing = 0;
    /* calc_h.F90:105-118 skips columns strictly inside a box by testing the
       loop variable k, which holds kend + 1 from the previous column's h loop
       (undefined before the first column with a free surface; the first
       column lies on the block face, outside every box): the test never
       holds for a box inside the block, so every column is computed */
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const size_t l0 = IX(b, i, j, 0);
        double zfs = 0.0;
        if (find_z_free_surface(z + l0, phi + l0, sk, nk, &zfs)) {
            for (int k = 0; k < nk; ++k) h[l0 + k * sk] = zfs - z[l0 + k * sk];
        } else {
            ++missing;
            for (int k = 0; k < nk; ++k) h[l0 + k * sk] = 0.0;
        }
    }
    /* calc_h.F90:195-345: per box, h extrapolated (4/3, -1/3) over whole
       columns onto the box's i faces, j faces and the four i-j corners, the
       face node and the node behind it; calc_h's own i_mysta = 1 and
       i_myend = ni - 2 (nodes without the block border) give the guards and
       clamp the face loops */
    const double i1 = 4.0 / 3.0, i2 = -1.0 / 3.0;
    const int ms = 1, mei = ni - 2, mej = nj - 2;
#define HCOL(dst, behind, a, c)                                                        \
    for (int k = 0; k < nk; ++k) {                                                     \
        const size_t o = (size_t)k * sk;                                               \
        h[(dst) + o] = i1 * h[(a) + o] + i2 * h[(c) + o];                              \
        h[(behind) + o] = h[(dst) + o];                                                \
    }
    for (int q = 0; q < b->nblk; ++q) {
        const int32_t *B = b->blank[q];
        const int ia = B[0], ib = B[1], ja = B[2], jb = B[3];
        const int *bt = &B[6];
        const int fi0 = bt[0] == 0 && ia > ms + 1, fi1 = bt[1] == 0 && ib < mei - 1;
        const int fj0 = bt[2] == 0 && ja > ms + 1, fj1 = bt[3] == 0 && jb < mej - 1;
        if (fi0)
            for (int j = imax(ms, ja); j <= imin(mej, jb); ++j)
                HCOL(IX(b, ia, j, 0), IX(b, ia + 1, j, 0), IX(b, ia - 1, j, 0), IX(b, ia - 2, j, 0))
        if (fi1)
            for (int j = imax(ms, ja); j <= imin(mej, jb); ++j)
                HCOL(IX(b, ib, j, 0), IX(b, ib - 1, j, 0), IX(b, ib + 1, j, 0), IX(b, ib + 2, j, 0))
        if (fj0)
            for (int i = imax(ms, ia); i <= imin(mei, ib); ++i)
                HCOL(IX(b, i, ja, 0), IX(b, i, ja + 1, 0), IX(b, i, ja - 1, 0), IX(b, i, ja - 2, 0))
        if (fj1)
            for (int i = imax(ms, ia); i <= imin(mei, ib); ++i)
                HCOL(IX(b, i, jb, 0), IX(b, i, jb - 1, 0), IX(b, i, jb + 1, 0), IX(b, i, jb + 2, 0))
        if (fi0 && fj0)
            HCOL(IX(b, ia, ja, 0), IX(b, ia + 1, ja + 1, 0), IX(b, ia - 1, ja - 1, 0), IX(b, ia - 2, ja - 2, 0))
        if (fi1 && fj0)
            HCOL(IX(b, ib, ja, 0), IX(b, ib - 1, ja + 1, 0), IX(b, ib + 1, ja - 1, 0), IX(b, ib + 2, ja - 2, 0))
        if (fi1 && fj1)
            HCOL(IX(b, ib, jb, 0), IX(b, ib - 1, jb - 1, 0), IX(b, ib + 1, jb + 1, 0), IX(b, ib + 2, jb + 2, 0))
        if (fi0 && fj1)
            HCOL(IX(b, ia, jb, 0), IX(b, ia + 1, jb - 1, 0), IX(b, ia - 1, jb + 1, 0), IX(b, ia - 2, jb + 2, 0))
    }
#undef HCOL
    return missing;
}

/* z_free_surface of every column as calc_h.F90:119-135 obtains it (0 where   */
/* find_z_free_surface finds no interval).  zfs: ni x nj, i fastest.          */
int oracle_free_surface(const OBlock *b, const double *z, const double *phi, double *zfs)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const size_t sk = (size_t)ni * nj;
    int missing = 0;
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const size_t l0 = IX(b, i, j, 0);
        double v = 0.0;
        if (!find_z_free_surface(z + l0, phi + l0, sk, nk, &v)) { ++missing; v = 0.0; }
        zfs[(size_t)i + (size_t)ni * j] = v;
    }
    return missing;
}

/* ------------------------------------------------------------------------- */
/* Geometric reinitialisation (hybrid_reinitialisation = .true., no global    */
/* mass correction): geometric_reinitialisation.F90:1-93 with                  */
/* GetTetrahedraList.F90, GetIsosurfaceAndDistances.F90,                       */
/* NeighboursPhiCorrection.F90 and TetrahedronMethods.F90.                     */
/* ------------------------------------------------------------------------- */

/* NORM2 as the Fortran runtime evaluates it for REAL(8): the largest |x| so
   far times sqrt(1 + sum of the others' squared ratios), rescaled when a
   larger element arrives (checked against amdflang on random vectors). */
static double fnorm2(const double v[3])
{
    double mx = 0.0, sum = 0.0;
    for (int q = 0; q < 3; ++q) {
        const double a = fabs(v[q]);
        if (mx == 0.0) {
            mx = a;
        } else if (a > mx) {
            const double t = mx / a, tsq = t * t;
            sum = sum * tsq;
            sum = sum + tsq;
            mx = a;
        } else {
            const double t = a / mx;
            sum = sum + t * t;
        }
    }
    return mx * sqrt(1.0 + sum);
}

static double fdot(const double a[3], const double b[3])
{
    double r = 0.0;
    for (int q = 0; q < 3; ++q) r = r + a[q] * b[q];
    return r;
}

static void fsub(const double a[3], const double b[3], double o[3])
{
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}

/* DistancePointLine, TetrahedronMethods.F90:1356-1399 */
static double dist_point_line(const double A[3], const double B[3], const double P[3])
{
    double AB[3], AP[3], BA[3], BP[3], c[3];
    fsub(B, A, AB); fsub(P, A, AP); fsub(A, B, BA); fsub(P, B, BP);
    int valid = 1;
    if (fnorm2(AB) < EPS_SIMS) valid = 0;
    if (fdot(AB, AP) < 0.0) valid = 0;
    if (fdot(BA, BP) < 0.0) valid = 0;
    if (valid) {
        cross3(AP, BP, c);
        return fnorm2(c) / fnorm2(AB);
    }
    const double da = fnorm2(AP), db = fnorm2(BP);
    return db < da ? db : da;
}

/* CheckProjectionPointOnATriangle3, TetrahedronMethods.F90:1593-1684:
   returns 1 and the distance to the projection when the projection falls
   strictly inside the triangle (the barycentric weights as the reference
   forms them) */
static int projection_inside(const double T[3][3], const double P[3], double *dist)
{
    double d12[3], d13[3], d23[3];
    fsub(T[0], T[1], d12); fsub(T[0], T[2], d13); fsub(T[1], T[2], d23);
    if (fnorm2(d12) <= EPS_SIMS || fnorm2(d13) <= EPS_SIMS || fnorm2(d23) <= EPS_SIMS) return 0;
    double u[3], v[3], w[3], N[3], c[3], pp[3], d[3];
    fsub(T[1], T[0], u); fsub(T[2], T[0], v); fsub(P, T[0], w);
    cross3(u, v, N);
    cross3(u, w, c);
    const double alpha = fdot(c, N) / fdot(N, N);
    cross3(w, v, c);
    const double beta = fdot(c, N) / fdot(N, N);
    const double gamma = (1.0 - beta) - alpha;
    for (int q = 0; q < 3; ++q) pp[q] = (alpha * T[0][q] + beta * T[1][q]) + gamma * T[2][q];
    fsub(P, pp, d);
    const double dpp = fnorm2(d);
    if (alpha > 0.0 && alpha < 1.0 && beta > 0.0 && beta < 1.0 && gamma > 0.0 && gamma < 1.0) {
        *dist = dpp;
        return 1;
    }
    return 0;
}

/* DistancePointTriangle, TetrahedronMethods.F90:1008-1070 */
static double dist_point_triangle(const double T[3][3], const double P[3])
{
    double d[3];
    double m[7];
    fsub(P, T[0], d); m[0] = fnorm2(d);
    fsub(P, T[1], d); m[1] = fnorm2(d);
    fsub(P, T[2], d); m[2] = fnorm2(d);
    m[3] = dist_point_line(T[0], T[1], P);
    m[4] = dist_point_line(T[1], T[2], P);
    m[5] = dist_point_line(T[2], T[0], P);
    double pd;
    const int n = projection_inside(T, P, &pd) ? 7 : 6;
    if (n == 7) m[6] = pd;
    double r = m[0];
    for (int q = 1; q < n; ++q) if (m[q] < r) r = m[q];
    return r;
}

/* GetTriangleArea, TetrahedronMethods.F90:691-707 */
static double triangle_area(const double T[3][3])
{
    double a[3], b[3], c[3];
    fsub(T[1], T[0], a); fsub(T[2], T[0], b);
    cross3(a, b, c);
    return 0.5 * fnorm2(c);
}

/* triangles of each case, TetrahedronMethods.F90:81-98 (edge indices) */
static const int TET_TRIANGLES[16][6] = {
    {-1, -1, -1, -1, -1, -1}, {0, 3, 2, -1, -1, -1}, {0, 1, 4, -1, -1, -1}, {1, 4, 2, 2, 4, 3},
    {1, 2, 5, -1, -1, -1}, {0, 3, 5, 0, 5, 1}, {0, 2, 5, 0, 5, 4}, {5, 4, 3, -1, -1, -1},
    {3, 4, 5, -1, -1, -1}, {4, 5, 0, 5, 2, 0}, {1, 5, 0, 5, 3, 0}, {5, 2, 1, -1, -1, -1},
    {3, 4, 2, 2, 4, 1}, {4, 1, 0, -1, -1, -1}, {2, 3, 0, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};

/* isosurface of MarchingTetrahedron, TetrahedronMethods.F90:26-664: the
   number of triangles (negative when the zero set is a vertex or an edge),
   their vertices and their total area */
static void marching_tet_isosurface(const double v[4][3], const double ph[4], double tri[2][3][3], int *ntri,
                                    double *area)
{
    const double tiny = 2.2250738585072014e-308;   /* tiny(eps_sims) */
    int sg[4], npos = 0, nneg = 0, nzero = 0;
    *ntri = 0; *area = 0.0;
    memset(tri, 0, sizeof(double) * 18);
    for (int q = 0; q < 4; ++q) {
        sg[q] = 0;
        if (ph[q] > tiny) sg[q] = 1;
        if (ph[q] < -tiny) sg[q] = -1;
        npos += sg[q] == 1; nneg += sg[q] == -1; nzero += sg[q] == 0;
    }
    if (npos == 4 || nneg == 4) return;
    int flag = 0;
    if (nzero > 0 && (nneg == 0 || npos == 0)) {
        if (nzero == 1) { *ntri = -1; return; }
        if (nzero == 2) { *ntri = -2; return; }
        if (nzero == 3) {
            /* the face of the three zero vertices, in vertex order */
            int r = 0;
            for (int q = 0; q < 4; ++q)
                if (sg[q] == 0) { memcpy(tri[0][r], v[q], sizeof(double) * 3); ++r; }
            *ntri = 1;
            *area = triangle_area(tri[0]);
            return;
        }
        return;                                         /* four zero vertices */
    }
    for (int q = 0; q < 4; ++q) if (sg[q] == -1) flag |= 1 << q;
    const int ef = TET_EDGE_FLAGS[flag];
    if (ef == 0) return;
    double E[6][3];
    for (int e = 0; e < 6; ++e) { E[e][0] = E[e][1] = E[e][2] = -9.9; }
    for (int e = 0; e < 6; ++e) {
        if (!(ef & (1 << e))) continue;
        const int s0 = TET_EDGE[e][0], s1 = TET_EDGE[e][1];
        const double den = ph[s1] - ph[s0];
        double sv = 0.0;
        if (!(fabs(den) < EPS_SIMS * EPS_SIMS * EPS_SIMS)) sv = ph[s1] / den;
        for (int c = 0; c < 3; ++c) E[e][c] = v[s1][c] + sv * (v[s0][c] - v[s1][c]);
    }
    for (int t = 0; t < 2; ++t) {
        if (TET_TRIANGLES[flag][3 * t] < 0) break;
        *ntri += 1;
        for (int q = 0; q < 3; ++q) memcpy(tri[t][q], E[TET_TRIANGLES[flag][3 * t + q]], sizeof(double) * 3);
        *area = *area + triangle_area(tri[t]);
    }
}

/* CellPnodeTetrahedronIdentification, TetrahedronMethods.F90:2124-2361
   (TriangulationCase 1): bits 1..6 = tetrahedra with a phase change, bits
   0..7 of *pnodes = their cell vertices */
static int cell_ktets(const double ph[8], int *pnodes)
{
    *pnodes = 0;
    int all_pos = 1, all_neg = 1;
    for (int q = 0; q < 8; ++q) { if (!(ph[q] > EPS_SIMS)) all_pos = 0; if (!(ph[q] < -EPS_SIMS)) all_neg = 0; }
    if (all_pos || all_neg) return 0;
    int pcell = 0;
    for (int q = 0; q < 8; ++q) if (fabs(ph[q]) < EPS_SIMS) pcell = 1;
    if (!pcell) {
        const double sgn = fsign1(ph[0]);
        for (int q = 1; q < 8; ++q) if (sgn * ph[q] < 0.0) { pcell = 1; break; }
    }
    if (!pcell) return 0;
    int tets = 0;
    for (int t = 0; t < 6; ++t) {
        int pct = 0, second = 1;
        double pa = ph[CELL_TETS[t][0]];
        if (pa <= EPS_SIMS) second = 0;
        if (fabs(pa) < EPS_SIMS) pct = 1;
        const double sgn = fsign1(pa);
        for (int q = 1; q < 4; ++q) {
            pa = ph[CELL_TETS[t][q]];
            if (sgn * pa < 0.0 || fabs(pa) < EPS_SIMS) {
                if (pa <= EPS_SIMS) second = 0;
                pct = 1;
            }
        }
        if (pct && !second) {
            tets |= 1 << (t + 1);
            for (int q = 0; q < 4; ++q) *pnodes |= 1 << CELL_TETS[t][q];
        }
    }
    return tets;
}

/* cell vertex q -> node offsets, GetTetrahedraList.F90:240-247 */
static const int CELL_VTX[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

typedef struct {
    int ntri;
    double area;
    double tri[2][3][3];
} OKTet;

/* every K-tetrahedron with a vertex at node m, in the order
   GetTetrahedraList.F90:346-418 appends them to the node's list (cells in
   k, j, i order, tetrahedra 1..6) */
static int node_ktets(const OBlock *b, const int *cell_tets, const int *cell_slot, int im, int jm, int km,
                      int out[24])
{
    int n = 0;
    for (int dk = -1; dk <= 0; ++dk)
    for (int dj = -1; dj <= 0; ++dj)
    for (int di = -1; di <= 0; ++di) {
        const int ci = im + di, cj = jm + dj, ck = km + dk;
        if (ci < 0 || cj < 0 || ck < 0 || ci > b->ni - 2 || cj > b->nj - 2 || ck > b->nk - 2) continue;
        const size_t c = (size_t)ci + (size_t)(b->ni - 1) * ((size_t)cj + (size_t)(b->nj - 1) * ck);
        const int tm = cell_tets[c];
        if (!tm) continue;
        int vq = -1;
        for (int q = 0; q < 8; ++q)
            if (CELL_VTX[q][0] == -di && CELL_VTX[q][1] == -dj && CELL_VTX[q][2] == -dk) vq = q;
        for (int t = 0; t < 6; ++t) {
            if (!(tm & (1 << (t + 1)))) continue;
            int has = 0;
            for (int q = 0; q < 4; ++q) if (CELL_TETS[t][q] == vq) has = 1;
            if (has) out[n++] = cell_slot[c] * 6 + t;
        }
    }
    return n;
}

/* the running minimum of GetIsosurfaceAndDistances.F90:201-233 over the
   isosurface of one K-tetrahedron */
static double ktet_min(const OKTet *K, const double P[3], double D)
{
    if (K->ntri > 0 && K->area > 0.0) {
        double Da = dist_point_triangle(K->tri[0], P);
        double ex = (fsign1(D - Da) + 1.0) / 2.0;
        D = D + ex * (Da - D);
        if (K->ntri > 1) {
            Da = dist_point_triangle(K->tri[1], P);
            ex = (fsign1(D - Da) + 1.0) / 2.0;
            D = D + ex * (Da - D);
        }
    }
    return D;
}

/* geometric_reinitialisation (GlobalMassCorrection = .false.):
   pc (phicorrected, in/out) receives the signed distance at the interface
   nodes and min(|pc|, distance) with the sign of pc at the nodes within 4
   nodes of them; ids (may be NULL) = 1 interface node, -1 such a neighbour,
   0 otherwise.  Returns the number of interface nodes. */
int oracle_geometric_reinit(const OBlock *b, const double *phi0, double *pc, int32_t *ids)
{
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const size_t n = (size_t)ni * nj * nk, nc = (size_t)(ni - 1) * (nj - 1) * (nk - 1);
    int *cell_tets = (int *)calloc(nc, sizeof(int));
    int *cell_slot = (int *)malloc(nc * sizeof(int));
    int *id = (int *)calloc(n, sizeof(int));
    int nslots = 0;
    for (int k = 0; k < nk - 1; ++k)
    for (int j = 0; j < nj - 1; ++j)
    for (int i = 0; i < ni - 1; ++i) {
        double ph[8];
        for (int q = 0; q < 8; ++q) ph[q] = phi0[IX(b, i + CELL_VTX[q][0], j + CELL_VTX[q][1], k + CELL_VTX[q][2])];
        int pn;
        const size_t c = (size_t)i + (size_t)(ni - 1) * ((size_t)j + (size_t)(nj - 1) * k);
        /* blanking: cells whose corner node lies in [ia, ib) x [ja, jb) x [ka, kb)
           of a box are not searched (GetTetrahedraList.F90:209-227) */
        int blanked = 0;
        for (int q = 0; q < b->nblk; ++q) {
            const int32_t *B = b->blank[q];
            if (i >= B[0] && i < B[1] && j >= B[2] && j < B[3] && k >= B[4] && k < B[5]) blanked = 1;
        }
        if (blanked) { cell_tets[c] = 0; cell_slot[c] = -1; continue; }
        cell_tets[c] = cell_ktets(ph, &pn);
        cell_slot[c] = cell_tets[c] ? nslots++ : -1;
        for (int q = 0; q < 8; ++q)
            if (pn & (1 << q)) id[IX(b, i + CELL_VTX[q][0], j + CELL_VTX[q][1], k + CELL_VTX[q][2])] = 1;
    }
    /* isosurfaces of the K-tetrahedra, GetIsosurfaceAndDistances.F90:78-130 */
    OKTet *kt = (OKTet *)calloc((size_t)nslots * 6 + 1, sizeof(OKTet));
    const double *X[3] = {b->x, b->y, b->z};
    for (int k = 0; k < nk - 1; ++k)
    for (int j = 0; j < nj - 1; ++j)
    for (int i = 0; i < ni - 1; ++i) {
        const size_t c = (size_t)i + (size_t)(ni - 1) * ((size_t)j + (size_t)(nj - 1) * k);
        if (!cell_tets[c]) continue;
        for (int t = 0; t < 6; ++t) {
            if (!(cell_tets[c] & (1 << (t + 1)))) continue;
            double v[4][3], ph[4];
            for (int q = 0; q < 4; ++q) {
                const int vq = CELL_TETS[t][q];
                const size_t l = IX(b, i + CELL_VTX[vq][0], j + CELL_VTX[vq][1], k + CELL_VTX[vq][2]);
                for (int a = 0; a < 3; ++a) v[q][a] = X[a][l];
                ph[q] = phi0[l];
            }
            OKTet *K = &kt[cell_slot[c] * 6 + t];
            marching_tet_isosurface(v, ph, K->tri, &K->ntri, &K->area);
        }
    }
    /* signed distances of the interface nodes, GetIsosurfaceAndDistances.F90:134-243,
       and their neighbours (-1) */
    double *sd = (double *)malloc(n * sizeof(double));
    int nint = 0;
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const size_t l = IX(b, i, j, k);
        if (id[l] <= 0) continue;
        ++nint;
        const double P[3] = {b->x[l], b->y[l], b->z[l]};
        double D = 10.0, SD = 99.9;
        for (int ks = (k - 4 < 0 ? 0 : k - 4); ks <= (k + 4 > nk - 1 ? nk - 1 : k + 4); ++ks)
        for (int js = (j - 4 < 0 ? 0 : j - 4); js <= (j + 4 > nj - 1 ? nj - 1 : j + 4); ++js)
        for (int is = (i - 4 < 0 ? 0 : i - 4); is <= (i + 4 > ni - 1 ? ni - 1 : i + 4); ++is) {
            const size_t ls = IX(b, is, js, ks);
            if (id[ls] > 0) {
                int lst[24];
                const int nkt = node_ktets(b, cell_tets, cell_slot, is, js, ks, lst);
                for (int q = 0; q < nkt; ++q) {
                    D = ktet_min(&kt[lst[q]], P, D);
                    const double ex2 = (fsign1(fabs(SD) - D) + 1.0) / 2.0;
                    SD = fabs(SD) + ex2 * (D - fabs(SD));
                    SD = fsign1(phi0[l]) * SD;
                }
            } else {
                id[ls] = -1;
            }
        }
        sd[l] = SD;
    }
    /* NeighboursPhiCorrection.F90:89-130 (C = 0): the interface nodes */
    for (size_t l = 0; l < n; ++l)
        if (id[l] > 0) pc[l] = sd[l] + 0.0;
    /* NeighboursPhiCorrection.F90:133-249: the neighbours */
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i) {
        const size_t l = IX(b, i, j, k);
        if (id[l] >= 0) continue;
        const int k0 = k - 4 < 0 ? 0 : k - 4, k1 = k + 4 > nk - 1 ? nk - 1 : k + 4;
        const int j0 = j - 4 < 0 ? 0 : j - 4, j1 = j + 4 > nj - 1 ? nj - 1 : j + 4;
        const int i0 = i - 4 < 0 ? 0 : i - 4, i1 = i + 4 > ni - 1 ? ni - 1 : i + 4;
        const double P[3] = {b->x[l], b->y[l], b->z[l]};
        for (int ks = k0; ks <= k1; ++ks)
        for (int js = j0; js <= j1; ++js)
        for (int is = i0; is <= i1; ++is) {
            if (id[IX(b, is, js, ks)] <= 0) continue;
            double D = 10.0;
            int lst[24];
            const int nkt = node_ktets(b, cell_tets, cell_slot, is, js, ks, lst);
            for (int q = 0; q < nkt; ++q) {
                D = ktet_min(&kt[lst[q]], P, D);
                const double ex2 = (fsign1(fabs(pc[l]) - D) + 1.0) / 2.0;
                const double aux = fabs(pc[l]) + ex2 * (D - fabs(pc[l]));
                pc[l] = fsign1(pc[l]) * aux;
            }
        }
    }
    if (ids) for (size_t l = 0; l < n; ++l) ids[l] = id[l] > 0 ? 1 : id[l];
    free(cell_tets); free(cell_slot); free(id); free(kt); free(sd);
    return nint;
}

/* ------------------------------------------------------------------------- */
/* R1: levelsetmethod, levelsetmethod.F90:135-326 (non-hybrid Sussman path,   */
/* no global mass correction).  phi: in = phi^n, out = phi^(n+1).  Returns    */
/* the pre-advection water volume (the value WriteTotalWaterVolume appends).  */
/* ------------------------------------------------------------------------- */
double oracle_levelset_step_h(const OBlock *b, const OReinit *p, double *phi,
                              const double *u, const double *v, const double *w,
                              int numiter_lsm, double delti_lsm, int call_levelset,
                              int call_reinit, int rn_freq, int ntime, int hybrid)
{
    const size_t n = (size_t)b->ni * b->nj * b->nk;
    double vol[3];
    int32_t *mask = (int32_t *)malloc(n * sizeof(int32_t));
    double *pc = (double *)calloc(n, sizeof(double));
    oracle_water_volume(b, phi, vol, mask);
    const double total = (vol[0] + vol[1]) + vol[2];
    /* the AdvectionNodes band restricts the advection only on the hybrid
       path (levelsetmethod.F90:198-199) */
    if (!call_reinit || !hybrid)
        for (size_t l = 0; l < n; ++l) mask[l] = 1;
    if (call_levelset) {
        double *U1 = (double *)malloc(n * sizeof(double));
        double *U2 = (double *)malloc(n * sizeof(double));
        double *U3 = (double *)malloc(n * sizeof(double));
        oracle_contra(b, u, v, w, U1, U2, U3);
        memcpy(pc, phi, n * sizeof(double));
        for (int it = 0; it < numiter_lsm; ++it) {
            oracle_rk3(b, phi, pc, mask, U1, U2, U3, delti_lsm);
            memcpy(pc, phi, n * sizeof(double));
        }
        free(U1); free(U2); free(U3);
    }
    if (call_reinit && rn_freq > 0 && ntime % rn_freq == 0) {
        if (hybrid) {
            /* levelsetmethod.F90:300-308 */
            for (size_t l = 0; l < n; ++l) {
                if (phi[l] < -EPS_SIMS) pc[l] = -9.9;
                if (phi[l] > EPS_SIMS) pc[l] = 9.9;
            }
            oracle_geometric_reinit(b, phi, pc, NULL);
        } else {
            oracle_reinit_eno2(b, p, phi, pc);
        }
        oracle_bcond(b, pc);
        memcpy(phi, pc, n * sizeof(double));
    }
    free(mask); free(pc);
    return total;
}

double oracle_levelset_step(const OBlock *b, const OReinit *p, double *phi,
                            const double *u, const double *v, const double *w,
                            int numiter_lsm, double delti_lsm, int call_levelset,
                            int call_reinit, int rn_freq, int ntime)
{
    return oracle_levelset_step_h(b, p, phi, u, v, w, numiter_lsm, delti_lsm, call_levelset, call_reinit,
                                  rn_freq, ntime, 0);
}
