// orderAD = 0: the first-order upwind right-hand side of the level-set
// advection (calc_RH_AD.F90:1-273) and the SSP-RK3 step built on it
// (ad_RKTVD3.F90:55-86 with calc_RH_AD, ad_RKTVD3.F90:57).  Unlike the WENO3
// path, the right-hand side is nonzero on the physical faces too (one-sided
// differences there), so every node with AdvectionNodes > 0 is advected and
// bcond_lsm then overwrites the faces.
#include "lsm_internal.hpp"

namespace lsm {
namespace {

struct UpArgs {
    const double *src, *pn;
    double *dst;
    const double *U1, *U2, *U3, *aj;
    const int32_t *mask;
    int ni, nj, nk;
    long sj, sk;
    double dd[3];
    double c1, c2, c3, dt;
};

// calc_RH_AD.F90:27-270 at node (i,j,k): per direction u+ = u/2 + |u/2|,
// u- = u/2 - |u/2| and dd (u+ (phi - phi_-1) + u- (phi_+1 - phi)), or
// (dd U)(phi_+1 - phi) on the low face and (dd U)(phi - phi_-1) on the high
// face; the xi term is assigned, the eta and zeta terms are added in front
// (rightH = term + rightH)
__device__ __forceinline__ double upwind_rhs_node(const UpArgs &a, const double *__restrict__ f, int i, int j,
                                                  int k, long l)
{
    const int idx[3] = {i, j, k};
    const int n[3] = {a.ni, a.nj, a.nk};
    const long st[3] = {1, a.sj, a.sk};
    const double *U[3] = {a.U1, a.U2, a.U3};
    const double fc = f[l];
    double r = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long s = st[d];
        const double ud = U[d][l];
        double t;
        if (idx[d] == 0) {
            t = (a.dd[d] * ud) * (f[l + s] - fc);
        } else if (idx[d] == n[d] - 1) {
            t = (a.dd[d] * ud) * (fc - f[l - s]);
        } else {
            const double u = 0.5 * ud;
            const double up = u + fabs(u), um = u - fabs(u);
            t = a.dd[d] * (up * (fc - f[l - s]) + um * (f[l + s] - fc));
        }
        r = (d == 0) ? t : t + r;
    }
    return r;
}

__device__ __forceinline__ bool node_of(const UpArgs &a, int &i, int &j, int &k, long &l)
{
    i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    k = (int)blockIdx.z;
    if (i >= a.ni || j >= a.nj) return false;
    l = (long)i + a.sj * j + a.sk * k;
    return true;
}

// calc_RH_AD: rh where AdvectionNodes > 0, other entries untouched
__global__ __launch_bounds__(256) void upwind_rhs_kernel(UpArgs a)
{
    int i, j, k;
    long l;
    if (!node_of(a, i, j, k, l)) return;
    if (a.mask && !(a.mask[l] > 0)) return;
    a.dst[l] = upwind_rhs_node(a, a.src, i, j, k, l);
}

// one RK stage: dst = ((c1 phi^n) + (c2 src)) - (((c3 aj) dt) rh(src)) where
// AdvectionNodes > 0, dst = src elsewhere
__global__ __launch_bounds__(256) void upwind_stage_kernel(UpArgs a)
{
    int i, j, k;
    long l;
    if (!node_of(a, i, j, k, l)) return;
    const double v = a.src[l];
    if (a.mask && !(a.mask[l] > 0)) { a.dst[l] = v; return; }
    const double rh = upwind_rhs_node(a, a.src, i, j, k, l);
    a.dst[l] = (a.c1 * a.pn[l] + a.c2 * v) - ((a.c3 * a.aj[l]) * a.dt) * rh;
}

UpArgs up_args(const Blk &b, const double *U1, const double *U2, const double *U3, const int32_t *mask)
{
    UpArgs a{};
    a.U1 = U1; a.U2 = U2; a.U3 = U3; a.aj = b.aj; a.mask = mask;
    a.ni = b.e.ni; a.nj = b.e.nj; a.nk = b.e.nk; a.sj = b.e.sj; a.sk = b.e.sk;
    a.dd[0] = b.dc; a.dd[1] = b.de; a.dd[2] = b.dz;
    return a;
}

dim3 up_grid(const Blk &b) { return dim3((unsigned)((b.e.ni + 63) / 64), (unsigned)((b.e.nj + 3) / 4), (unsigned)b.e.nk); }

}  // namespace

int launch_upwind_rhs(const Blk &b, const double *phi, const double *U1, const double *U2, const double *U3,
                      const int32_t *mask, double *rhs, hipStream_t s)
{
    if (b.e.nk > 65535) return fail(LSM_ERR_ARG, "nk > 65535");
    UpArgs a = up_args(b, U1, U2, U3, mask);
    a.src = phi; a.dst = rhs;
    upwind_rhs_kernel<<<up_grid(b), dim3(64, 4, 1), 0, s>>>(a);
    return check_hip(hipGetLastError(), "upwind_rhs_kernel");
}

int launch_upwind_rk3(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                      const double *U2, const double *U3, double dt, const double *const src[3],
                      double *const dst[3], hipStream_t s)
{
    if (b.e.nk > 65535) return fail(LSM_ERR_ARG, "nk > 65535");
    static const double C[3][3] = {{1.0, 0.0, 1.0},
                                   {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    UpArgs a = up_args(b, U1, U2, U3, mask);
    a.pn = phi_n; a.dt = dt;
    for (int st = 0; st < 3; ++st) {
        a.src = src[st]; a.dst = dst[st];
        a.c1 = C[st][0]; a.c2 = C[st][1]; a.c3 = C[st][2];
        {
            KTimer kt(LSM_KT_ADVECT_STAGE, s);
            upwind_stage_kernel<<<up_grid(b), dim3(64, 4, 1), 0, s>>>(a);
        }
        int rc = check_hip(hipGetLastError(), "upwind_stage_kernel");
        if (rc) return rc;
        if ((rc = launch_bcond(b, dst[st], s))) return rc;
    }
    return LSM_OK;
}

}  // namespace lsm
