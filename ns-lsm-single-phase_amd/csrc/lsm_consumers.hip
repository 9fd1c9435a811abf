// Caller-side consumers of phi, run once per time step before the flow solver
// (mg_driver.F90:106-150): the ENO2 gradient of phi (calc_phi_gradient_ENO2),
// the local water depth h = z_fs - z of every column (calc_h with
// find_z_free_surface) and the horizontal ENO2 gradient of h
// (calc_h_gradient_ENO2).  One block, six physical faces.
#include "lsm_advect_tile.hpp"
#include "lsm_reinit_common.hpp"

namespace lsm {
namespace {

using rn::Geo;
using rn::make_geo;

constexpr int GX = 64, GY = 4;   // gradient tile: 64 i-lanes x 4 j-rows of one k-plane

// calc_phi_gradient_ENO2.F90:106-2020 (NC = 3) and calc_h_gradient_ENO2.F90
// (NC = 2: the x and y components): the calc_NormGrad_d stencils with the
// node's own value as the ENO2 upwind sign, combined with the metrics of the
// node.  Interior, face and edge nodes are written; the eight vertices are
// not (the routines have no vertex blocks and their output is intent(inout)).
// One thread per node; the stencil neighbours come through L1/L2, and the
// tiles of consecutive planes go to one XCD so its L2 holds the j/k halos.
// Blanking (interior nodes only, calc_phi_gradient_ENO2.F90:133-158,
// 216-231): the stencils one node off a box's i or j face are biased away from
// it and the xi (eta) derivative is zero on the box's i (j) faces, at every k.
template <int NC>
__global__ __launch_bounds__(GX *GY) void gradient_eno2_kernel(Geo g, const double *__restrict__ f,
                                                               int order_bc, int eno_bc, int tiles_i,
                                                               int tiles_j, int ntiles,
                                                               double *__restrict__ o0,
                                                               double *__restrict__ o1,
                                                               double *__restrict__ o2, Blank bl)
{
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    if (t >= ntiles) return;
    const int ti = t % tiles_i;
    const int r = t / tiles_i;
    const int tj = r % tiles_j;
    const int k = r / tiles_j;
    const int i = ti * GX + (int)threadIdx.x;
    const int j = tj * GY + (int)threadIdx.y;
    if (i >= g.ni || j >= g.nj) return;
    bool on[3];
    on[0] = (i == 0 || i == g.ni - 1);
    on[1] = (j == 0 || j == g.nj - 1);
    on[2] = (k == 0 || k == g.nk - 1);
    if (on[0] && on[1] && on[2]) return;
    const long l = (long)i + g.sj * j + g.sk * k;
    double gr[3];
    int blo = 0, bhi = 0;
    bool z0 = false, z1 = false;
    if (bl.n && !on[0] && !on[1] && !on[2]) {
        bool mb = false, mf = false, ml = false, mr = false;
        bl.bias(i, j, mb, mf, ml, mr);
        blo = (mb ? 1 : 0) | (ml ? 2 : 0);
        bhi = (mf ? 1 : 0) | (mr ? 2 : 0);
        for (int q = 0; q < bl.n; ++q) {
            const int *B = bl.b[q];
            const bool jin = j >= B[2] && j <= B[3], iin = i >= B[0] && i <= B[1];
            if ((i == B[0] || i == B[1]) && jin) z0 = true;
            if ((j == B[2] || j == B[3]) && iin) z1 = true;
        }
    }
    rn::curv_derivs(g, f, f[l], order_bc, eno_bc != 0, i, j, k, l, on, gr, blo, bhi);
    if (z0) gr[0] = 0.0;
    if (z1) gr[1] = 0.0;
    o0[l] = (g.m[0][l] * gr[0] + g.m[3][l] * gr[1]) + g.m[6][l] * gr[2];
    o1[l] = (g.m[1][l] * gr[0] + g.m[4][l] * gr[1]) + g.m[7][l] * gr[2];
    if (NC == 3) o2[l] = (g.m[2][l] * gr[0] + g.m[5][l] * gr[1]) + g.m[8][l] * gr[2];
}

constexpr int HX = 64;           // free-surface search: 64 columns (consecutive i) per workgroup
constexpr int HS = 8;            // k-slices per column
constexpr int kNoHit = 0x7fffffff;

struct Column {
    int i, j, ts;
    bool col;                    // (i, j) inside the block
    long l0;                     // node (i, j, 0)
};

__device__ __forceinline__ Column column_of(int ni, int nj, long sj)
{
    Column c;
    const int tiles_i = (ni + HX - 1) / HX;
    c.i = ((int)blockIdx.x % tiles_i) * HX + (int)threadIdx.x;
    c.j = (int)blockIdx.x / tiles_i;
    c.ts = (int)threadIdx.y;
    c.col = (c.i < ni) && (c.j < nj);
    c.l0 = (long)c.i + sj * c.j;
    return c;
}

// find_z_free_surface.F90:16-58 for the workgroup's 64 columns: the first
// pair (k, k+1) from the bottom with phi(k) phi(k+1) <= 0.  Each of the 8
// k-slices reads every pair of its part of the column (so the loads do not
// wait on the search) and keeps its first hit; the lowest hit over the slices
// is the reference's.  Returns kNoHit for a column without a sign change.
// Every thread of the column gets the result.
__device__ __forceinline__ int first_crossing(const Column &c, int nk, long sk, const double *__restrict__ phi,
                                              int (*kf_s)[HX])
{
    const int np = nk - 1;
    const int cp = (np + HS - 1) / HS;
    const int k0 = c.ts * cp, k1 = min(k0 + cp, np);
    int kf = kNoHit;
    if (c.col && k0 < k1) {
        double pa = phi[c.l0 + sk * k0];
        for (int k = k0; k < k1; ++k) {
            const double pb = phi[c.l0 + sk * (k + 1)];
            if (kf == kNoHit && pa * pb <= 0.0) kf = k;
            pa = pb;
        }
    }
    kf_s[c.ts][threadIdx.x] = kf;
    __syncthreads();
    int kmin = kNoHit;
#pragma unroll
    for (int s = 0; s < HS; ++s) kmin = min(kmin, kf_s[s][threadIdx.x]);
    return kmin;
}

// the four cases of find_z_free_surface.F90:22-54 in the interval (k, k+1)
__device__ __forceinline__ double crossing_height(double p0, double p1, double z0, double z1)
{
    if (fabs(p0) < kEpsSims && fabs(p1) > kEpsSims) return z0;
    if (fabs(p1) < kEpsSims && fabs(p0) > kEpsSims) return z1;
    if (fabs(p1) < kEpsSims && fabs(p0) < kEpsSims) return 0.5 * (z1 + z0);
    return z0 + ((z1 - z0) * (-p0)) / (p1 - p0);
}

// calc_h.F90:101-185: z_fs of every column (zfs, ni x nj, may be NULL; 0 in
// a column without a free surface, the value find_z_free_surface leaves in
// z_free_surface) and h = z_fs - z (h may be NULL; 0 in a column without a
// free surface).  `missing` (may be NULL) counts the columns without a free
// surface, where the reference prints 'free surface not found'.
__global__ __launch_bounds__(HX *HS) void free_surface_kernel(int ni, int nj, int nk, long sj, long sk,
                                                              const double *__restrict__ z,
                                                              const double *__restrict__ phi,
                                                              double *__restrict__ h,
                                                              double *__restrict__ zfs_out,
                                                              int32_t *__restrict__ missing)
{
    __shared__ int kf_s[HS][HX];
    const Column c = column_of(ni, nj, sj);
    const int kf = first_crossing(c, nk, sk, phi, kf_s);
    const bool found = c.col && kf != kNoHit;
    double zfs = 0.0;
    if (found)
        zfs = crossing_height(phi[c.l0 + sk * kf], phi[c.l0 + sk * (kf + 1)], z[c.l0 + sk * kf],
                              z[c.l0 + sk * (kf + 1)]);
    if (c.ts == 0) {
        if (zfs_out && c.col) zfs_out[(long)c.i + (long)ni * c.j] = zfs;
        if (missing) {
            const unsigned long long lost = __ballot(c.col && !found);
            if (threadIdx.x == 0 && lost) atomicAdd(missing, (int32_t)__popcll(lost));
        }
    }
    if (!h || !c.col) return;
    const int ck = (nk + HS - 1) / HS;
    const int ka = c.ts * ck, kb = min(ka + ck, nk);
    for (int k = ka; k < kb; ++k) {
        const long l = c.l0 + sk * k;
        h[l] = found ? zfs - z[l] : 0.0;
    }
}

// Reverse mode of the free-surface height (no counterpart in the reference):
// adj_phi = (d z_fs / d phi)^T adj_zfs, written at every node.  Only the
// interpolating case depends on phi: z_fs = z0 + (z1 - z0)(-p0)/(p1 - p0)
// gives d/dp0 = -(z1 - z0) p1 / (p1 - p0)^2, d/dp1 = (z1 - z0) p0 / (p1 - p0)^2;
// a surface at a node (|phi| < eps) and a column without one contribute 0.
__global__ __launch_bounds__(HX *HS) void free_surface_adjoint_kernel(int ni, int nj, int nk, long sj, long sk,
                                                                      const double *__restrict__ z,
                                                                      const double *__restrict__ phi,
                                                                      const double *__restrict__ adj_zfs,
                                                                      double *__restrict__ adj_phi)
{
    __shared__ int kf_s[HS][HX];
    const Column c = column_of(ni, nj, sj);
    const int kf = first_crossing(c, nk, sk, phi, kf_s);
    if (!c.col) return;
    double a0 = 0.0, a1 = 0.0;
    if (kf != kNoHit) {
        const double p0 = phi[c.l0 + sk * kf], p1 = phi[c.l0 + sk * (kf + 1)];
        const bool interp = !((fabs(p0) < kEpsSims && fabs(p1) > kEpsSims) ||
                              (fabs(p1) < kEpsSims && fabs(p0) > kEpsSims) ||
                              (fabs(p1) < kEpsSims && fabs(p0) < kEpsSims));
        if (interp) {
            const double dz = z[c.l0 + sk * (kf + 1)] - z[c.l0 + sk * kf];
            const double den = (p1 - p0) * (p1 - p0);
            const double lam = adj_zfs[(long)c.i + (long)ni * c.j];
            a0 = lam * (-(dz * p1) / den);
            a1 = lam * ((dz * p0) / den);
        }
    }
    const int ck = (nk + HS - 1) / HS;
    const int ka = c.ts * ck, kb = min(ka + ck, nk);
    for (int k = ka; k < kb; ++k) adj_phi[c.l0 + sk * k] = (k == kf) ? a0 : (k == kf + 1) ? a1 : 0.0;
}

__global__ void zero_counter(int32_t *c) { *c = 0; }

}  // namespace

int launch_gradient_eno2(const Blk &b, const double *f, int nc, int order_bc, int eno_bc, double *o0,
                         double *o1, double *o2, hipStream_t s)
{
    const int tiles_i = (b.e.ni + GX - 1) / GX, tiles_j = (b.e.nj + GY - 1) / GY;
    const long ntiles = (long)tiles_i * tiles_j * b.e.nk;
    if (ntiles > (1L << 30)) return fail(LSM_ERR_ARG, "block too large for the gradient kernel");
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
    const dim3 blk(GX, GY);
    if (nc == 3)
        gradient_eno2_kernel<3><<<grid, blk, 0, s>>>(make_geo(b), f, order_bc, eno_bc, tiles_i, tiles_j,
                                                     (int)ntiles, o0, o1, o2, b.bl);
    else
        gradient_eno2_kernel<2><<<grid, blk, 0, s>>>(make_geo(b), f, order_bc, eno_bc, tiles_i, tiles_j,
                                                     (int)ntiles, o0, o1, nullptr, b.bl);
    return check_hip(hipGetLastError(), "gradient_eno2_kernel");
}

int launch_free_surface(const Blk &b, const double *z, const double *phi, double *h, double *zfs,
                        int32_t *missing, hipStream_t s)
{
    const long nwg = (long)((b.e.ni + HX - 1) / HX) * b.e.nj;
    if (nwg > (1L << 30)) return fail(LSM_ERR_ARG, "block too large for the free-surface search");
    if (missing) {
        zero_counter<<<1, 1, 0, s>>>(missing);
        int rc = check_hip(hipGetLastError(), "zero_counter");
        if (rc) return rc;
    }
    free_surface_kernel<<<(unsigned)nwg, dim3(HX, HS), 0, s>>>(b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, z, phi,
                                                               h, zfs, missing);
    int rc = check_hip(hipGetLastError(), "free_surface_kernel");
    if (rc || !h || !b.bl.n) return rc;
    // calc_h.F90:195-345: h extrapolated onto the blanking boxes' i and j faces
    // over whole columns.  (The column skip of calc_h.F90:105-118 tests the
    // loop variable k, which holds kend + 1 from the previous column, so it
    // never holds for a box inside the block: every column is computed.)
    return launch_blank_extrap(b, h, 1, false, false, true, s);
}

int launch_free_surface_adjoint(const Blk &b, const double *z, const double *phi, const double *adj_zfs,
                                double *adj_phi, hipStream_t s)
{
    const long nwg = (long)((b.e.ni + HX - 1) / HX) * b.e.nj;
    if (nwg > (1L << 30)) return fail(LSM_ERR_ARG, "block too large for the free-surface search");
    free_surface_adjoint_kernel<<<(unsigned)nwg, dim3(HX, HS), 0, s>>>(b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk,
                                                                       z, phi, adj_zfs, adj_phi);
    return check_hip(hipGetLastError(), "free_surface_adjoint_kernel");
}

}  // namespace lsm
