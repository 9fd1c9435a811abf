// Device-side arithmetic of the level-set transport path.
//
// Every function keeps the operation order of the reference routine it
// restates (cited per function); the library is compiled with
// -ffp-contract=off so no multiply-add is fused behind the source's back.
// The deliberate exceptions are weno3_fast() / weno3_dir_fast() / quot(),
// which rewrite the WENO3 weights with fewer divisions and allow fused
// multiply-adds (selected by LSM_ADVECT_FAST; LSM_ADVECT_EXACT keeps the
// reference arithmetic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsm {

// workgroup -> tile, keeping consecutive tiles on one XCD: the dispatcher
// deals workgroups round-robin over the 8 XCDs, so workgroup b runs on XCD
// b % 8; giving each XCD a contiguous run of tiles keeps the j/k halo reuse
// inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}


constexpr double kEpsSims = 2.220446049250313e-16;   // epsilon(1._rdf), precision.F90:62
constexpr double kPi = 3.141592653589793;            // precision.F90:51
constexpr double kTiny = 2.2250738585072014e-308;    // tiny(eps_sims), TetrahedronMethods.F90:177

__device__ __forceinline__ double sign1(double x) { return copysign(1.0, x); }

// GetWENO3Reconstruction, AdvectionMethods.F90:195-245 (reference order: four divisions)
__device__ __forceinline__ double weno3_exact(double f0, double f1, double f2, double f3, double ws)
{
    const double d0 = 2.0 / 3.0, d1 = 1.0 / 3.0, e = 0.000001;
    const bool pos = ws > 0.0;
    const double fL = pos ? f0 : f3;
    const double fC = pos ? f1 : f2;
    const double fR = pos ? f2 : f1;
    const double b0 = (fC - fR) * (fC - fR);
    const double b1 = (fL - fC) * (fL - fC);
    const double a0 = d0 / ((b0 + e) * (b0 + e));
    const double a1 = d1 / ((b1 + e) * (b1 + e));
    const double w0 = a0 / (a0 + a1);
    const double w1 = a1 / (a0 + a1);
    const double u0 = 0.5 * fC + 0.5 * fR;
    const double u1 = (-0.5) * fL + 1.5 * fC;
    return u0 * w0 + u1 * w1;
}

// The same reconstruction with the weights folded into one quotient:
//   w0 = d0/(b0+e)^2 / (d0/(b0+e)^2 + d1/(b1+e)^2) = 2 A1 / (2 A1 + A0),
//   A0 = (b0+e)^2, A1 = (b1+e)^2, so the face value is
//   ((fC+fR) A1 + (1.5 fC - 0.5 fL) A0) / (2 A1 + A0).
// Agrees with weno3_exact to a few ulp of the face value.
__device__ __forceinline__ double weno3_fast(double f0, double f1, double f2, double f3, double ws)
{
#pragma clang fp contract(fast)
    const bool pos = ws > 0.0;
    const double fL = pos ? f0 : f3;
    const double fC = pos ? f1 : f2;
    const double fR = pos ? f2 : f1;
    const double e = 0.000001;
    const double s0 = (fC - fR) * (fC - fR) + e;
    const double s1 = (fL - fC) * (fL - fC) + e;
    const double A0 = s0 * s0;
    const double A1 = s1 * s1;
    const double u1 = 1.5 * fC - 0.5 * fL;
    const double num = (fC + fR) * A1 + u1 * A0;
    const double den = 2.0 * A1 + A0;
    return num / den;
}

template <bool FAST>
__device__ __forceinline__ double weno3(double f0, double f1, double f2, double f3, double ws)
{
    if constexpr (FAST) return weno3_fast(f0, f1, f2, f3, ws);
    else return weno3_exact(f0, f1, f2, f3, ws);
}

// n / d through the hardware reciprocal: two Newton steps and one residual
// correction (accurate to about an ulp; d is a sum of squares in (1e-24, 1e4)
// here, far from the range where the IEEE division's scaling matters)
__device__ __forceinline__ double quot(double n, double d)
{
#pragma clang fp contract(fast)
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    double q = n * r;
    const double res = fma(-d, q, n);
    return fma(r, res, q);
}

// One direction of calc_RH_AD_WENO3.F90:195-254: U/J * (d * (Phi+ - Phi-)),
// with the outward face forced upwind at the first/last interior node
// (calc_RH_AD_WENO3.F90:201-237).  Reference arithmetic.
__device__ __forceinline__ double weno3_dir_exact(double m2, double m1, double c0, double p1, double p2,
                                                  double ws, bool forced_minus, bool forced_plus, double d)
{
    const double wsp = forced_plus ? 1.0 : ws;
    const double f3 = forced_plus ? 0.0 : p2;
    const double fp = weno3_exact(m1, c0, p1, f3, wsp);
    const double wsm = forced_minus ? -1.0 : ws;
    const double f0 = forced_minus ? 0.0 : m2;
    const double fm = weno3_exact(f0, m1, c0, p1, wsm);
    return ws * (d * (fp - fm));
}

// The same direction term with the weights of weno3_fast, sharing work
// between the two faces: with Dk the squared differences of consecutive
// stencil values, Sk = (Dk + 1e-6)^2 serve as A0/A1 of both faces and both
// upwind choices (4 squarings instead of 8), and the two face quotients are
// combined, Phi+ - Phi- = (Np Dm - Nm Dp) / (Dp Dm), into one division.
__device__ __forceinline__ double weno3_dir_fast(double m2, double m1, double c0, double p1, double p2,
                                                 double ws, bool forced_minus, bool forced_plus, double d)
{
#pragma clang fp contract(fast)
    const double e = 0.000001;
    const double am = m2 - m1, a0 = m1 - c0, a1 = c0 - p1, a2 = p1 - p2;
    double t;
    t = am * am + e; const double Sm = t * t;
    t = a0 * a0 + e; const double S0 = t * t;
    t = a1 * a1 + e; const double S1 = t * t;
    t = a2 * a2 + e; const double S2 = t * t;
    const bool pp = forced_plus || ws > 0.0;      // plus face: left-biased stencil
    const bool pm = !forced_minus && ws > 0.0;    // minus face: left-biased stencil
    const double Ap = pp ? S0 : S2;
    const double up = pp ? (1.5 * c0 - 0.5 * m1) : (1.5 * p1 - 0.5 * p2);
    const double Np = (c0 + p1) * Ap + up * S1;
    const double Dp = 2.0 * Ap + S1;
    const double Am = pm ? Sm : S1;
    const double um = pm ? (1.5 * m1 - 0.5 * m2) : (1.5 * c0 - 0.5 * p1);
    const double Nm = (m1 + c0) * Am + um * S0;
    const double Dm = 2.0 * Am + S0;
    return (ws * d) * quot(Np * Dm - Nm * Dp, Dp * Dm);
}

template <bool FAST>
__device__ __forceinline__ double weno3_dir(double m2, double m1, double c0, double p1, double p2,
                                            double ws, bool forced_minus, bool forced_plus, double d)
{
    if constexpr (FAST) return weno3_dir_fast(m2, m1, c0, p1, p2, ws, forced_minus, forced_plus, d);
    else return weno3_dir_exact(m2, m1, c0, p1, p2, ws, forced_minus, forced_plus, d);
}

// GetENO2Reconstruction, AdvectionMethods.F90:64-191
__device__ __forceinline__ double eno2(double f0, double fLL, double fL, double fC, double fR,
                                       double fRR, bool left_out, bool right_out)
{
    double a, b, Mf, dp, dm;
    a = (fR - 2.0 * fC) + fL;
    if (right_out) {
        dp = (fR - fC) - 0.5 * a;
    } else {
        b = (fRR - 2.0 * fR) + fC;
        Mf = (fabs(b) < fabs(a)) ? b : a;
        dp = (fR - fC) - 0.5 * Mf;
    }
    if (left_out) {
        dm = (fC - fL) + 0.5 * a;
    } else {
        b = (fC - 2.0 * fL) + fLL;
        Mf = (fabs(b) < fabs(a)) ? b : a;
        dm = (fC - fL) + 0.5 * Mf;
    }
    // a left-outbounded stencil takes precedence (the reference tests it first)
    if (left_out && right_out) {
        b = (fRR - 2.0 * fR) + fC;
        Mf = (fabs(b) < fabs(a)) ? b : a;
        dp = (fR - fC) - 0.5 * Mf;
        dm = (fC - fL) + 0.5 * a;
    }
    const double s = (fabs(f0) > kEpsSims) ? sign1(f0) : 0.0;
    if (s * (fR - fC) < 0.0 && s * (fC - fL) < -(s * (fR - fC))) return dp;
    if (s * (fC - fL) > 0.0 && s * (fR - fC) > -(s * (fC - fL))) return dm;
    return 0.5 * (dp + dm);
}

// ENO2Boundary, AdvectionMethods.F90:654-707
__device__ __forceinline__ double eno2_boundary(double f0, double f1, double f2, int dir)
{
    const double as = (double)dir;
    const double D1 = as * (f1 - f0);
    const double D2 = as * (f2 - f1);
    const double D3 = as * (f2 - f0) / 2.0;
    if (fabs(D1 - D2) < kEpsSims) return D1;
    if (fabs(D1) < fabs(D2)) return (2.0 * D1 + D2) / 3.0;
    if (fabs(D2) < fabs(D1)) return (D1 + 2.0 * D2) / 3.0;
    if (fabs(D1) == fabs(D2)) return (4.0 * D1 + D2) / 5.0;
    if (fabs(D1) < fabs(D3) && fabs(D2) < fabs(D3)) return ((3.0 * D1 + 6.0 * D2) - D3) / 7.0;
    if (fabs(D1) >= fabs(D3) && fabs(D2) >= fabs(D3)) return D3;
    return D1;
}

// BiasedDerivative, AdvectionMethods.F90:248-302
__device__ __forceinline__ double biased(double f0, double f1, double f2, double f3, int order,
                                         int dir, bool enobc)
{
    const double as = (double)dir;
    double r = 0.0;
    if (order == 1) r = as * (-(1.0 * f0) + 1.0 * f1);
    if (order == 2) {
        r = as * ((-(3.0 * f0) + 4.0 * f1) - 1.0 * f2) / 2.0;
        if (enobc) r = eno2_boundary(f0, f1, f2, dir);
    }
    if (order == 3) r = as * (((-(11.0 * f0) + 18.0 * f1) - 9.0 * f2) + 2.0 * f3) / 6.0;
    return r;
}

// ---- linear forms of the ENO2 / biased derivatives on the branch taken ----
// For the adjoint: each derivative above is, once its comparisons are fixed,
// a linear combination of the stencil values.  The *_branch functions return
// a small code for the branch taken, evaluating every comparison with the
// same expressions as the forward functions; the *_coef functions turn a code
// into the coefficients of the combination.

// branch of eno2(): bits 0-1 the result (0 dp, 1 dm, 2 their mean), bit 2 the
// plus-side second difference (fC, fR, fRR) instead of (fL, fC, fR), bit 3 the
// minus-side one (fLL, fL, fC) instead of (fL, fC, fR)
__device__ __forceinline__ int eno2_branch(double f0, double fLL, double fL, double fC, double fR,
                                           double fRR, bool left_out, bool right_out)
{
    const double a = (fR - 2.0 * fC) + fL;
    int br = 0;
    if (!right_out || left_out) {
        const double b = (fRR - 2.0 * fR) + fC;
        if (fabs(b) < fabs(a)) br |= 4;
    }
    if (!left_out) {
        const double b = (fC - 2.0 * fL) + fLL;
        if (fabs(b) < fabs(a)) br |= 8;
    }
    const double s = (fabs(f0) > kEpsSims) ? sign1(f0) : 0.0;
    if (s * (fR - fC) < 0.0 && s * (fC - fL) < -(s * (fR - fC))) br |= 0;
    else if (s * (fC - fL) > 0.0 && s * (fR - fC) > -(s * (fC - fL))) br |= 1;
    else br |= 2;
    return br;
}

// coefficients of eno2() on branch br for (fLL, fL, fC, fR, fRR)
__device__ __forceinline__ void eno2_coef(int br, double e[5])
{
    // dp = (fR - fC) - 0.5 Mf+,  dm = (fC - fL) + 0.5 Mf-
    double dp[5] = {0.0, 0.0, -1.0, 1.0, 0.0};
    double dm[5] = {0.0, -1.0, 1.0, 0.0, 0.0};
    if (br & 4) { dp[2] -= 0.5; dp[3] += 1.0; dp[4] -= 0.5; }
    else { dp[1] -= 0.5; dp[2] += 1.0; dp[3] -= 0.5; }
    if (br & 8) { dm[0] += 0.5; dm[1] -= 1.0; dm[2] += 0.5; }
    else { dm[1] += 0.5; dm[2] -= 1.0; dm[3] += 0.5; }
    const int pick = br & 3;
#pragma unroll
    for (int q = 0; q < 5; ++q) e[q] = pick == 0 ? dp[q] : (pick == 1 ? dm[q] : 0.5 * (dp[q] + dm[q]));
}

// branch of eno2_boundary(), 0..6 in the order of its tests
__device__ __forceinline__ int eno2_boundary_branch(double f0, double f1, double f2, int dir)
{
    const double as = (double)dir;
    const double D1 = as * (f1 - f0);
    const double D2 = as * (f2 - f1);
    const double D3 = as * (f2 - f0) / 2.0;
    if (fabs(D1 - D2) < kEpsSims) return 0;
    if (fabs(D1) < fabs(D2)) return 1;
    if (fabs(D2) < fabs(D1)) return 2;
    if (fabs(D1) == fabs(D2)) return 3;
    if (fabs(D1) < fabs(D3) && fabs(D2) < fabs(D3)) return 4;
    if (fabs(D1) >= fabs(D3) && fabs(D2) >= fabs(D3)) return 5;
    return 6;
}

// coefficients of biased() for (f0, f1, f2, f3); br: eno2_boundary_branch
// (used for order 2 with enobc only)
__device__ __forceinline__ void biased_coef(int br, int order, int dir, bool enobc, double e[4])
{
    const double as = (double)dir;
    e[0] = e[1] = e[2] = e[3] = 0.0;
    if (order == 1) { e[0] = -as; e[1] = as; }
    if (order == 2) {
        if (enobc) {
            // D1 = as (-1, 1, 0), D2 = as (0, -1, 1), D3 = as (-1/2, 0, 1/2)
            double w1 = 1.0, w2 = 0.0, w3 = 0.0;
            if (br == 1) { w1 = 2.0 / 3.0; w2 = 1.0 / 3.0; }
            else if (br == 2) { w1 = 1.0 / 3.0; w2 = 2.0 / 3.0; }
            else if (br == 3) { w1 = 4.0 / 5.0; w2 = 1.0 / 5.0; }
            else if (br == 4) { w1 = 3.0 / 7.0; w2 = 6.0 / 7.0; w3 = -1.0 / 7.0; }
            else if (br == 5) { w1 = 0.0; w3 = 1.0; }
            e[0] = as * (-w1 - 0.5 * w3);
            e[1] = as * (w1 - w2);
            e[2] = as * (w2 + 0.5 * w3);
        } else {
            e[0] = -1.5 * as; e[1] = 2.0 * as; e[2] = -0.5 * as;
        }
    }
    if (order == 3) { e[0] = -11.0 / 6.0 * as; e[1] = 3.0 * as; e[2] = -1.5 * as; e[3] = 1.0 / 3.0 * as; }
}

// The library sin / cos, called out of line: they are needed only inside the
// interface band (|phi| <= eps), and inlined their registers count against
// every node of the stencil kernels that use them (same functions, same
// results)
static __device__ __noinline__ double band_sin(double x) { return sin(x); }
static __device__ __noinline__ double band_cos(double x) { return cos(x); }

// d sign_phi0 / d phi0
__device__ __forceinline__ double sign_phi0_deriv(double p, double eps)
{
    if (fabs(p) > eps) return 0.0;
    return (1.0 - band_cos((kPi * p) / eps)) / eps;
}

// GetSignPhi0, AdvectionMethods.F90:576-587
__device__ __forceinline__ double sign_phi0(double p, double eps)
{
    if (fabs(p) > eps) return sign1(p);
    return p / eps - (1.0 / kPi) * band_sin((kPi * p) / eps);
}

// deltaf, global_lsm.F90:120-133
__device__ __forceinline__ double deltaf(double v, double eps)
{
    if (fabs(v) > eps) return 0.0;
    return (1.0 / (2.0 * eps)) * (1.0 + band_cos((kPi * v) / eps));
}

// GetLambdaMassCorrection, global_lsm.F90:200-219, given del = deltaf(d)
__device__ __forceinline__ double lambda_mass_del(double del, double s, double g)
{
    const double num = -((del * s) * (1.0 - g));
    const double den = (del * del) * g;
    if (fabs(den) < kEpsSims) return 0.0;
    return num / den;
}

// The update of one node of a Sussman stage (reinitialisation_test2.F90:137-181):
// (c1 d^n + c2 d) - (c3 dtau) rhs, rhs = s (|grad d| - 1) - lambda deltaf(d) |grad d|.
// s = sign_phi0(phi0) and del = deltaf(d) are the caller's (each evaluated once
// per node: the band sin / cos are out-of-line calls the compiler cannot merge).
__device__ __forceinline__ double sussman_update(double c1, double c2, double c3dt, double ngv, double dv,
                                                 double dnv, double s, double del)
{
    double rhs = s * (ngv - 1.0);
    const double lam = lambda_mass_del(del, s, ngv);
    const double aux = del * ngv;
    rhs = rhs - lam * aux;
    return (c1 * dnv + c2 * dv) - c3dt * rhs;
}

// One face of bcond_lsm.F90:53-287.  bt: 0 copy the neighbour, 1 second-order
// extrapolation switching to first order where |phi(2)| >= BigPhi, 2 nothing
// except linear extrapolation on the im face.  `self` is the node's current
// value (returned unchanged for type 2 on the other faces).
__device__ __forceinline__ double bc_face(int bt, double p1, double p2, double self, double big,
                                          bool is_im)
{
    if (bt == 0) return p1;
    if (bt == 1) {
        const double a2 = 4.0 / 3.0, b2 = -1.0 / 3.0, a1 = 1.0, b1 = 0.0;
        const double ex = (1.0 + sign1(fabs(p2) - big)) / 2.0;
        const double ta = a2 + ex * (a1 - a2);
        const double tb = b2 + ex * (b1 - b2);
        return ta * p1 + tb * p2;
    }
    if (bt == 2 && is_im) return 2.0 * p1 - 1.0 * p2;
    return self;
}

// block-level extents passed to kernels (0-based local indices)
struct Ext {
    int ni, nj, nk;      // array extents
    long sj, sk;         // strides of j and k
    int faces;           // LSM_FACE_* bits
    __host__ __device__ long idx(int i, int j, int k) const { return (long)i + sj * j + sk * k; }
};

// Blanking regions (global_app.F90:30-43) passed to kernels by value: n
// boxes, each (ia, ib, ja, jb, ka, kb) as 0-based node indices followed by
// blktype(1:4).  The loops run over at most four boxes of kernel-argument
// memory, uniform across the wavefront.
struct Blank {
    int n;
    int b[4][10];        // LSM_MAX_BLANK boxes
    // the node lies in a box (faces included)
    __host__ __device__ bool inside(int i, int j, int k) const
    {
        for (int q = 0; q < n; ++q)
            if (i >= b[q][0] && i <= b[q][1] && j >= b[q][2] && j <= b[q][3] && k >= b[q][4] && k <= b[q][5])
                return true;
        return false;
    }
    // the node lies strictly inside a box
    __host__ __device__ bool interior(int i, int j, int k) const
    {
        for (int q = 0; q < n; ++q)
            if (i > b[q][0] && i < b[q][1] && j > b[q][2] && j < b[q][3] && k > b[q][4] && k < b[q][5])
                return true;
        return false;
    }
    // the cell with lower corner (i, j, k) has that corner in [ia, ib) x [ja, jb) x [ka, kb)
    __host__ __device__ bool cell(int i, int j, int k) const
    {
        for (int q = 0; q < n; ++q)
            if (i >= b[q][0] && i < b[q][1] && j >= b[q][2] && j < b[q][3] && k >= b[q][4] && k < b[q][5])
                return true;
        return false;
    }
    // WENO3 stencils biased away from the i and j faces of a box next to the
    // node, at every k (calc_RH_AD_WENO3.F90:152-188)
    __host__ __device__ void bias(int i, int j, bool &mb, bool &mf, bool &ml, bool &mr) const
    {
        for (int q = 0; q < n; ++q) {
            const bool jin = j >= b[q][2] && j <= b[q][3], iin = i >= b[q][0] && i <= b[q][1];
            if (i == b[q][0] - 1 && jin) mf = true;
            if (i == b[q][1] + 1 && jin) mb = true;
            if (j == b[q][2] - 1 && iin) mr = true;
            if (j == b[q][3] + 1 && iin) ml = true;
        }
    }
};

}  // namespace lsm
