// Host-side plumbing shared by the kernel files: the validated block
// description, error reporting and launcher declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lsm_device.hpp"
#include "lsm_hip.h"

namespace lsm {

static_assert(LSM_MAX_BLANK == 4, "Blank holds four boxes");

// Everything a kernel needs about one block, by value.
struct Blk {
    Ext e;
    int btype[6];
    double dc, de, dz, big_phi;
    const double *csi0, *csi1, *csi2, *eta0, *eta1, *eta2, *zet0, *zet1, *zet2;
    const double *aj, *x, *y, *z;
    long n;                 // nodes in the block
    Blank bl;               // blanking regions, 0-based (bl.n = 0: none)
    const double *gten;     // the caller's metric tensor (6 x n), or NULL
};

// thread-local last-error text
int fail(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);

// validates the descriptor; `need_metrics` / `need_coords` select which
// device pointers must be present.  Blanking regions (nblk != 0) are
// rejected unless `blanking` is kBlankUse (the entry point implements them)
// or kBlankIgnore (its reference routine does not read them)
enum BlankPolicy { kBlankReject = 0, kBlankUse = 1, kBlankIgnore = 2 };
int make_blk(const LsmBlock *b, Blk *out, bool need_metrics, bool need_coords,
             BlankPolicy blanking = kBlankReject);

// library-owned buffers per (device, stream, purpose)
enum { kBufWork = 0, kBufStore = 1, kBufMulti = 2 };
int cached_buffer(int purpose, size_t bytes, hipStream_t s, void **ptr);
int default_work(size_t bytes, hipStream_t s, void **ptr);

// carve `n` bytes, 256-byte aligned, from a workspace cursor
struct Carver {
    char *base;
    size_t size, used;
    void *take(size_t n)
    {
        size_t off = (used + 255) & ~(size_t)255;
        if (off + n > size) return nullptr;
        used = off + n;
        return base + off;
    }
};

// per-kernel launch timing (lsm_timing_*): begin/end bracket one launch
struct KTimer {
    int id;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(int kid, hipStream_t st);
    ~KTimer();
};

// whether lsm_timing_enable(1) is in force (no event recording may enter a graph capture)
bool timing_on();

// ---- launchers (lsm_advect.hip) ----
int launch_contra(const Blk &b, const double *u, const double *v, const double *w, long stride,
                  double *U1, double *U2, double *U3, hipStream_t s);
int launch_weno3_rhs(const Blk &b, const double *phi, const double *U1, const double *U2,
                     const double *U3, const int32_t *mask, double *rhs, int mode, hipStream_t s);
// bcond_lsm: the physical faces, then the blanking regions
int launch_bcond(const Blk &b, double *phi, hipStream_t s);
// the blanking part of bcond_lsm alone (nothing when b.bl.n == 0)
int launch_bcond_blanking(const Blk &b, double *phi, hipStream_t s);
// the box-face extrapolations of bcond_lsm (m = 0, zero = switch_order = true,
// full_k = false) and of calc_h (m = 1, fixed second order, whole columns)
int launch_blank_extrap(const Blk &b, double *f, int m, bool zero, bool switch_order, bool full_k,
                        hipStream_t s);
int launch_rk3(const Blk &b, double *phi_np1, const double *phi_n, const int32_t *mask,
               const double *U1, const double *U2, const double *U3, double dt, int mode,
               double *scratch, hipStream_t s, bool folded = false);
// the three stage results of one RK3 step (phi^(1), phi^(2), phi^(3) = phi^(n+1))
int launch_rk3_stages(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                      const double *U2, const double *U3, double dt, int mode, double *out1,
                      double *out2, double *out3, hipStream_t s);
int launch_fold(const Blk &b, const double *U1, const double *U2, const double *U3, double dt,
                double *V1, double *V2, double *V3, int *bad, hipStream_t s);
int launch_copy(const double *src, double *dst, long n, hipStream_t s);

// ---- launchers (lsm_geometric.hip): hybrid reinitialisation ----
size_t geometric_work(const Blk &b);
// pc = -9.9 where phi < -eps, 9.9 where phi > eps, base (0 when NULL) elsewhere
int launch_hybrid_init(const double *phi, const double *base, double *pc, long n, hipStream_t s);
int launch_geometric(const Blk &b, const double *phi0, double *pc, int32_t *ids, Carver &c, hipStream_t s);

// ---- launchers (lsm_upwind.hip): orderAD = 0 ----
int launch_upwind_rhs(const Blk &b, const double *phi, const double *U1, const double *U2, const double *U3,
                      const int32_t *mask, double *rhs, hipStream_t s);
int launch_upwind_rk3(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                      const double *U2, const double *U3, double dt, const double *const src[3],
                      double *const dst[3], hipStream_t s);

// ---- launchers (lsm_consumers.hip) ----
// nc = 3: calc_phi_gradient_ENO2, nc = 2: calc_h_gradient_ENO2 (o2 unused)
int launch_gradient_eno2(const Blk &b, const double *f, int nc, int order_bc, int eno_bc, double *o0,
                         double *o1, double *o2, hipStream_t s);
// calc_h (h may be NULL) and the per-column free-surface height (zfs may be NULL)
int launch_free_surface(const Blk &b, const double *z, const double *phi, double *h, double *zfs,
                        int32_t *missing, hipStream_t s);
int launch_free_surface_adjoint(const Blk &b, const double *z, const double *phi, const double *adj_zfs,
                                double *adj_phi, hipStream_t s);

// ---- launchers (lsm_adjoint.hip) ----
size_t rk3_adjoint_work(const Blk &b);
// `stages` (optional): the three stage results of the forward step, recorded
// earlier; otherwise they are recomputed into the workspace
int launch_rk3_adjoint(const Blk &b, const double *phi_n, const int32_t *mask, const double *U1,
                       const double *U2, const double *U3, double dt, int mode, const double *adj_np1,
                       double *adj_n, double *adjU1, double *adjU2, double *adjU3, Carver &w,
                       hipStream_t s, const double *const *stages = nullptr, bool adj_in_place = false);

// ---- launchers (lsm_reinit.hip) ----
size_t reinit_eno2_work(const Blk &b);
// the reinitialisation's metric tensor (6 x n doubles, component c at gt + c * n)
int launch_metric_tensor(const Blk &b, double *gt, hipStream_t s);
int launch_metric_tensor(const double *const m[9], long n, double *gt, hipStream_t s);
size_t reinit_godunov_work(const Blk &b);
size_t lambda_simpson_work(const Blk &b);
int launch_normgrad(const Blk &b, const double *d, const double *phi0, int order_bc, int eno_bc,
                    double *ng, hipStream_t s);
int launch_reinit_eno2(const Blk &b, const LsmReinitParams &p, const double *phi0, double *out,
                       int32_t *iters, Carver &w, hipStream_t s);
int launch_sign_ini(const Blk &b, const double *phi0, double eps, int m_sphi0, double *sgn,
                    double *ng, hipStream_t s);
int launch_rh_rn(const Blk &b, const double *d, const double *sgn, double *rh, hipStream_t s);
int launch_lambda_rn(const Blk &b, const double *phi0, const double *ng0, const double *d,
                     double dt, double eps, double *lam, hipStream_t s);
int launch_lambda_rn_simpson(const Blk &b, const double *phi0, const double *ng0, const double *d, double dt,
                             double eps, double *lam, Carver &w, hipStream_t s);
int launch_reinit_godunov(const Blk &b, const LsmReinitParams &p, const double *phi0,
                          double *out, Carver &w, hipStream_t s);

// The Sussman iterations of launch_reinit_eno2 with a tape for the adjoint:
// per iteration the update mask, per stage and node the linearisation of the
// update (direction weights q, branch codes, the s(phi0) term T inside the
// interface band); *iters (device) the number of iterations that ran.
constexpr int kMaxTraceIters = 64;
struct ReinitTape {
    int nit;
    uint8_t *conv[kMaxTraceIters];
    double *q[kMaxTraceIters][3][3];     // [iteration][stage][direction]
    uint16_t *code[kMaxTraceIters][3];
    double *T[kMaxTraceIters][3];
    int32_t *iters;
};
size_t reinit_trace_work(const Blk &b, int itermax);
// runs the iterations from phi0 into `out` (required, not aliasing phi0),
// recording the tape in w
int launch_reinit_eno2_trace(const Blk &b, const LsmReinitParams &p, const double *phi0, ReinitTape &t,
                             double *out, Carver &w, hipStream_t s);
// the tape's buffers carved from w as launch_reinit_eno2_trace carves them
// (no launch): re-opens a tape recorded earlier in the same memory
int carve_reinit_trace(const Blk &b, const LsmReinitParams &p, ReinitTape &t, Carver &w, void **ctl,
                       double **A, double **B);

// ---- launchers (lsm_reinit_adjoint.hip) ----
size_t reinit_eno2_adjoint_work(const Blk &b, int itermax);
// adj_phi0 = (d out / d phi0)^T adj_out for out = reinit(phi0); `out`
// (optional) receives the forward result.  `pre`: a trace of phi0 already
// recorded (then `out` is ignored and the trace's workspace stays in use)
int launch_reinit_eno2_adjoint(const Blk &b, const LsmReinitParams &p, const double *phi0,
                               const double *adj_out, double *adj_phi0, double *out, Carver &w,
                               hipStream_t s, const ReinitTape *pre = nullptr);
// bcond_lsm transposed in place on adj; phi is the forward field after bcond_lsm
int launch_bcond_adjoint(const Blk &b, const double *phi, double *adj, hipStream_t s);

// ---- launchers (lsm_volume.hip) ----
size_t water_volume_work(const Blk &b);
// timed = false: no launch timing events (the allocation-free entry point)
int launch_water_volume(const Blk &b, const double *phi, double *vol, int32_t *mask, Carver &w,
                        hipStream_t s, bool timed = true);

}  // namespace lsm
