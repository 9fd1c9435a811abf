/*
 * lsm_hip.h -- C ABI of the MI355X level-set interface-transport library.
 *
 * Each entry point replaces one routine of the reference level-set path
 * (jcsandov/NS-LSM-Single-Phase, src/), keeping its argument meaning.  The
 * reference routines are internal procedures of `levelsetmethod` and read
 * their grid, metrics and parameters by host association; here that state is
 * passed explicitly through `LsmBlock` and `LsmReinitParams`.
 *
 * Conventions
 *  - All arrays are device pointers to float64 (int32 for masks) with the
 *    reference's Fortran memory order: element (i,j,k) of a block with
 *    extents il:iu, jl:ju, kl:ku lives at
 *        (i-il) + (iu-il+1)*((j-jl) + (ju-jl+1)*(k-kl)).
 *  - Vector fields are structure-of-arrays (three separate arrays) where the
 *    reference stores them interleaved (csi(1:3,i,j,k)).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *    asynchronous on that stream; results are ready after the stream is
 *    synchronised.  Scalar results (volumes, iteration counts) are written to
 *    device memory.
 *  - Every call returns LSM_OK (0) or a negative LSM_ERR_* code; the message
 *    of the last error on the calling thread is in lsm_last_error().
 *  - `work` arguments are device scratch of at least lsm_workspace_bytes()
 *    bytes; NULL lets the library use its own buffer for the calling
 *    (device, stream) pair (allocated on first use, grown on demand, kept
 *    until exit).  Calls on different streams never share library-owned
 *    memory, so they may run concurrently; calls on one stream from several
 *    host threads at once must be serialised by the caller.
 */
#ifndef LSM_HIP_H
#define LSM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSM_OK                 0
#define LSM_ERR_ARG           -1   /* bad argument (NULL pointer, bad extent, bad option) */
#define LSM_ERR_UNSUPPORTED   -2   /* configuration outside what the library implements */
#define LSM_ERR_HIP           -3   /* a HIP runtime call failed */
#define LSM_ERR_WORKSPACE     -4   /* workspace too small */

/* physical-boundary face bits (the reference's my* == mpi_proc_null tests) */
#define LSM_FACE_I1  0x01   /* myback  : i = il+igp */
#define LSM_FACE_IM  0x02   /* myfront : i = iu-igp */
#define LSM_FACE_J1  0x04   /* myleft  : j = jl+jgp */
#define LSM_FACE_JM  0x08   /* myright : j = ju-jgp */
#define LSM_FACE_K1  0x10   /* mydown  : k = kl+kgp */
#define LSM_FACE_KM  0x20   /* myup    : k = ku-kgp */
#define LSM_FACE_ALL 0x3f

/* advection arithmetic */
#define LSM_ADVECT_FAST   0   /* WENO3 weights with one division per face value */
#define LSM_ADVECT_EXACT  1   /* the reference's operation order (four divisions) */
#define LSM_ADVECT_UPWIND1 2  /* orderAD = 0: first-order upwind right-hand side (calc_RH_AD.F90);
                                 lsm_rk3_step and lsm_levelset_step only */

/* workspace queries */
#define LSM_OP_RK3            1
#define LSM_OP_REINIT_ENO2    2
#define LSM_OP_REINIT_GODUNOV 3
#define LSM_OP_WATER_VOLUME   4
#define LSM_OP_LEVELSET_STEP  5
#define LSM_OP_RK3_ADJOINT    6
#define LSM_OP_REINIT_ENO2_ADJOINT 7   /* sized by lsm_adjoint_workspace_bytes */
#define LSM_OP_STEP_ADJOINT   8        /* sized by lsm_adjoint_workspace_bytes */
#define LSM_OP_STEP_TRACE     9        /* sized by lsm_adjoint_workspace_bytes */
#define LSM_OP_STEP_ADJOINT_RECORDED 10 /* sized by lsm_adjoint_workspace_bytes */
#define LSM_OP_REINIT_GEOMETRIC 11
#define LSM_OP_LAMBDA_RN_SIMPSON 12

/* One blanking region (global_app.F90:30-43, set up by init.F90:1216-1567;
 * with one block per rank li_blk_* = le_blk_*): a box of nodes given in the
 * block's array indices (li_blk_ia..li_blk_kb, inclusive) and blktype(1:4)
 * of its i and j faces (0: bcond_lsm extrapolates the face from outside).
 * The box must lie in the block, with ia <= iu-1 and ib >= il+1 (likewise
 * for j) so that bcond_lsm's writes next to each face stay in the array. */
#define LSM_MAX_BLANK 4
typedef struct LsmBlank {
    int32_t ia, ib, ja, jb, ka, kb;
    int32_t blktype[4];
} LsmBlank;

/* One curvilinear block (the host-associated state of levelsetmethod.F90:1-44).
 * Zero-initialise it: nblk = 0 means no blanking. */
typedef struct LsmBlock {
    int32_t il, iu, jl, ju, kl, ku;   /* array extents */
    int32_t igp, jgp, kgp;            /* ghost widths */
    int32_t phys_faces;               /* LSM_FACE_* bits of physical boundaries */
    int32_t btype[6];                 /* btype_lsm(1:6, myzone): 0 copy, 1 extrapolate, 2 advect */
    double dc, de, dz;                /* 1/d xi, 1/d eta, 1/d zeta (dc = de = dz = 1) */
    double big_phi;                   /* BigPhi of bcond_lsm */
    const double *csi[3];             /* xi_x, xi_y, xi_z */
    const double *eta[3];             /* eta_x, eta_y, eta_z */
    const double *zet[3];             /* zeta_x, zeta_y, zeta_z */
    const double *aj;                 /* aj = 1/det(dx/dxi) */
    const double *x, *y, *z;          /* node coordinates (water volume only) */
    int32_t nblk;                     /* blanking regions in blank[] (nblk of global_app.F90:31) */
    LsmBlank blank[LSM_MAX_BLANK];    /* read by lsm_weno3_rhs, lsm_bcond, lsm_rk3_step,
                                         lsm_water_volume(_noalloc), lsm_reinit_geometric,
                                         lsm_levelset_step, the gradients of phi and h,
                                         lsm_calc_h and the lsm_host_* adapter; the entry
                                         points whose reference routines do not read the
                                         blanking arrays ignore it; the adjoints, the folded
                                         step and lsm_free_surface reject nblk != 0 */
    const double *gten;               /* optional (NULL: none): the metric tensor of the
                                         Sussman reinitialisation, 6 x n doubles written by
                                         lsm_metric_tensor from csi, eta, zet; with it the
                                         reinitialisation stages read 6 fields per node
                                         instead of 9 and skip the per-call tensor pass.
                                         The caller keeps it in step with the metrics. */
} LsmBlock;

/* Parameters of the pseudo-time reinitialisation (global_lsm.F90:17-58). */
typedef struct LsmReinitParams {
    double epslsm;                    /* interface half width */
    double eps_reinit;                /* epsReinitialisation */
    double rn_timestep;               /* RNtimestep (d tau) */
    int32_t rn_itermax;               /* RNitermax */
    int32_t order_bc;                 /* OrderReinitialisationBoundaries (1, 2, 3) */
    int32_t eno_bc;                   /* ENOBCReinitialisation (0/1) */
    int32_t m_sphi0;                  /* Godunov path: 1 piecewise-sine sign, else phi/sqrt(phi^2+(|grad|eps)^2) */
    int32_t m_ti;                     /* Godunov path: 1 SSP-RK3, else forward Euler */
    int32_t hybrid;                   /* hybrid_reinitialisation: geometric instead of Sussman
                                         in lsm_levelset_step (GlobalMassCorrection = .false.) */
    int32_t lambda_method;            /* sussman_correction_method (Godunov path): 0 the pointwise
                                         lambda, 1 the filtered one (lsm_lambda_rn_simpson) */
} LsmReinitParams;

/* Library version string. */
const char *lsm_version(void);

/* Message of the last error returned on the calling thread. */
const char *lsm_last_error(void);

/* Bytes of device scratch an operation needs for this block (LSM_OP_*). */
size_t lsm_workspace_bytes(const LsmBlock *blk, int op);

/* Bytes of device scratch of the adjoints that depend on the number of
 * reinitialisation iterations (LSM_OP_REINIT_ENO2_ADJOINT, LSM_OP_STEP_ADJOINT;
 * both keep the stage inputs of every iteration).  0 for rn_itermax > 64. */
size_t lsm_adjoint_workspace_bytes(const LsmBlock *blk, int op, int rn_itermax);

/* rhs_contra_j.F90:2-50 (rhs_contra_j): ucn_j = (1/aj) * (grad xi^j . u) at
 * every owned node.  u, v, w are read with element stride `vel_stride`
 * (1 for separate arrays, 4 to read q(2:4,...) of the reference directly by
 * passing q+1, q+2, q+3). */
int lsm_contra_velocity(const LsmBlock *blk, const double *u, const double *v,
                        const double *w, int64_t vel_stride,
                        double *ucn1, double *ucn2, double *ucn3, void *stream);

/* calc_RH_AD_WENO3.F90:1-407 (calc_RH_AD_WENO3): right-hand side of the
 * WENO3 advection operator on interior nodes with adv_mask != 0 (adv_mask may
 * be NULL: every node).  Other nodes of `rhs` are left untouched (the caller
 * zeroes rhs, as ad_RKTVD3.F90:55 does). */
int lsm_weno3_rhs(const LsmBlock *blk, const double *phi, const double *ucn1,
                  const double *ucn2, const double *ucn3, const int32_t *adv_mask,
                  double *rhs, int mode, void *stream);

/* bcond_lsm.F90:2-568 (bcond_lsm): boundary values on the physical faces, in
 * place, in the reference's face order (i1, im, j1, jm, k1, km). */
int lsm_bcond(const LsmBlock *blk, double *phi, void *stream);

/* calc_RH_AD.F90:1-273 (calc_RH_AD, orderAD = 0): the first-order upwind
 * right-hand side (u/2 +- |u/2| split, one-sided differences times U on the
 * physical faces) at every node with AdvectionNodes > 0 (adv_mask may be
 * NULL: every node); other entries of rhs keep their values. */
int lsm_upwind_rhs(const LsmBlock *blk, const double *phi, const double *ucn1, const double *ucn2,
                   const double *ucn3, const int32_t *adv_mask, double *rhs, void *stream);

/* ad_RKTVD3.F90:1-88 (ad_RKTVD3): one SSP-RK3 step of the WENO3 advection,
 * phi_np1 = RK3(phi_n), with bcond_lsm after each stage.  phi_np1 and phi_n
 * must not alias.  adv_mask may be NULL (every node advected). */
int lsm_rk3_step(const LsmBlock *blk, double *phi_np1, const double *phi_n,
                 const int32_t *adv_mask, const double *ucn1, const double *ucn2,
                 const double *ucn3, double delti_lsm, int mode,
                 void *work, size_t work_bytes, void *stream);

/* Frozen-velocity form of ad_RKTVD3 (no reference counterpart; an
 * arithmetic rearrangement of ad_RKTVD3.F90:70-72 for callers that advance
 * many steps with one velocity, e.g. a frozen velocity field): the operator
 * V_d = (aj * delti_lsm * dc_d) * ucn_d is built once, then each stage
 * streams V (3 fields) instead of ucn and aj (4 fields).  Fast arithmetic
 * only.  lsm_advect_fold waits for the stream and fails with
 * LSM_ERR_UNSUPPORTED if aj <= 0 or aj = inf anywhere (the upwind sign of V would differ
 * from that of ucn) or delti_lsm <= 0. */
int lsm_advect_fold(const LsmBlock *blk, const double *ucn1, const double *ucn2,
                    const double *ucn3, double delti_lsm, double *v1, double *v2, double *v3,
                    void *stream);
int lsm_rk3_step_folded(const LsmBlock *blk, double *phi_np1, const double *phi_n,
                        const int32_t *adv_mask, const double *v1, const double *v2,
                        const double *v3, void *work, size_t work_bytes, void *stream);

/* Adjoint of lsm_rk3_step (reverse mode of ad_RKTVD3.F90:1-88 with
 * calc_RH_AD_WENO3.F90:1-407 and bcond_lsm.F90:2-568; the reference has no
 * adjoint): given adj_np1 = dJ/dphi^(n+1), writes adj_n = dJ/dphi^n and, when
 * adj_ucn1..3 are given, ADDS dJ/d(U/J) of this step to them (zero them before
 * the first step of a trajectory).  The forward stages are recomputed from
 * phi_n with the same arguments (mode selects their arithmetic).  The WENO3
 * weights are differentiated; the upwind choice, the forced boundary stencils
 * and the BigPhi switch of bcond_lsm are taken as in the forward.
 * Deterministic.  Workspace: lsm_workspace_bytes(blk, LSM_OP_RK3_ADJOINT)
 * (20 fields). */
int lsm_rk3_step_adjoint(const LsmBlock *blk, const double *phi_n, const int32_t *adv_mask,
                         const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm,
                         int mode, const double *adj_np1, double *adj_n, double *adj_ucn1,
                         double *adj_ucn2, double *adj_ucn3, void *work, size_t work_bytes,
                         void *stream);

/* Adjoint of lsm_bcond (reverse mode of bcond_lsm.F90:2-568; the reference
 * has no adjoint): adj <- (d bcond / d phi)^T adj in place, the face phases
 * transposed in reverse order (k, j, i); phi is the forward field AFTER
 * bcond_lsm (its BigPhi switch and boundary types are taken as in the
 * forward). */
int lsm_bcond_adjoint(const LsmBlock *blk, const double *phi, double *adj, void *stream);

/* calc_NormGrad_d.F90:1-2448 (calc_NormGrad_d): |grad d| with ENO2
 * differences inside and biased differences of order `order_bc` (optionally
 * ENO2Boundary when eno_bc != 0 and order_bc == 2) on faces, edges and
 * vertices; phizero selects the ENO2 upwind side. */
int lsm_normgrad_eno2(const LsmBlock *blk, const double *d, const double *phizero,
                      int order_bc, int eno_bc, double *normgrad, void *stream);

/* geometric_reinitialisation.F90:1-93 (hybrid_reinitialisation = .true.,
 * GlobalMassCorrection = .false.) with GetTetrahedraList.F90,
 * GetIsosurfaceAndDistances.F90 and NeighboursPhiCorrection.F90:
 * phicorrected (in/out) receives, at every vertex of a tetrahedron whose phi0
 * changes sign (an interface node), its signed distance to the marching-
 * tetrahedra isosurface of phi0 within 4 nodes, and at every other node within
 * 4 nodes of an interface node min(|phicorrected|, that distance) with the
 * sign of phicorrected; other nodes keep their values.  ids (device int32,
 * may be NULL) = 1 interface node, -1 such a neighbour, 0 other.  Needs the
 * coordinates x, y, z of blk.  Synchronises the stream once (the number of
 * cut cells sizes the isosurface store, which the library keeps per
 * (device, stream) pair).
 * Workspace: lsm_workspace_bytes(blk, LSM_OP_REINIT_GEOMETRIC). */
int lsm_reinit_geometric(const LsmBlock *blk, const double *phi0, double *phicorrected, int32_t *ids,
                         void *work, size_t work_bytes, void *stream);

/* calc_phi_gradient_ENO2.F90:1-2025 (calc_phi_gradient_ENO2, called every
 * step by mg_driver.F90:106-116): grad phi = (g1, g2, g3) with the
 * calc_NormGrad_d stencils, the node's own phi as the ENO2 upwind sign.
 * Written on interior, face and edge nodes; the eight vertices keep the
 * caller's values (the reference's output is intent(inout) and has no
 * vertex blocks).  Needs the csi/eta/zet metrics of blk only. */
int lsm_phi_gradient_eno2(const LsmBlock *blk, const double *phi, int order_bc, int eno_bc,
                          double *g1, double *g2, double *g3, void *stream);

/* calc_h_gradient_ENO2.F90:1-2018 (calc_h_gradient_ENO2,
 * mg_driver.F90:140-150): the x and y components of grad h, same stencils
 * and vertex rule as lsm_phi_gradient_eno2. */
int lsm_h_gradient_eno2(const LsmBlock *blk, const double *h, int order_bc, int eno_bc,
                        double *g1, double *g2, void *stream);

/* calc_h.F90:1-394 with find_z_free_surface.F90:1-62 (calc_h,
 * mg_driver.F90:127-135): in every (i,j) column, z_fs from the first k with
 * phi(k) phi(k+1) <= 0 (at a node where |phi| < eps, midway between two such
 * nodes, else linearly interpolated) and h = z_fs - z; h = 0 in a column
 * without a free surface.  `missing` (device int32, may be NULL) receives the
 * number of such columns (the reference prints 'free surface not found'
 * for each). */
int lsm_calc_h(const LsmBlock *blk, const double *z, const double *phi, double *h,
               int32_t *missing, void *stream);

/* find_z_free_surface.F90:1-62 over every (i,j) column, as calc_h.F90:101-135
 * calls it: zfs (device, ni x nj, i fastest) receives z_free_surface of each
 * column (0 in a column without a free surface, the value the reference
 * leaves there); missing as in lsm_calc_h.  The surface elevation of a
 * water column of depth d is zfs - d. */
int lsm_free_surface(const LsmBlock *blk, const double *z, const double *phi, double *zfs,
                     int32_t *missing, void *stream);

/* Reverse mode of lsm_free_surface (the reference has no adjoint): adj_phi
 * (written at every node) = (d zfs / d phi)^T adj_zfs.  Only the
 * interpolating case of find_z_free_surface depends on phi (a surface at a
 * node and a column without one contribute 0); the interval is the one the
 * forward finds. */
int lsm_free_surface_adjoint(const LsmBlock *blk, const double *z, const double *phi,
                             const double *adj_zfs, double *adj_phi, void *stream);

/* reinitialisation_test2.F90:1-218 (reinitialisation_test2, Sussman path,
 * non-hybrid): phi_out = reinitialised phi0.  iters_done (device int32, may be
 * NULL) receives the number of pseudo-time iterations that ran.  At nodes off
 * the block faces |grad d| is evaluated as sqrt(g^T G g) from the metric
 * tensor G_ab = grad xi_a . grad xi_b (lsm_metric_tensor) and the ENO2
 * derivatives g: the reference's |csi g_xi + eta g_eta + zet g_zeta|
 * (calc_NormGrad_d.F90:145-159) up to rounding. */
int lsm_reinit_sussman_eno2(const LsmBlock *blk, const LsmReinitParams *p,
                            const double *phi0, double *phi_out, int32_t *iters_done,
                            void *work, size_t work_bytes, void *stream);

/* No reference counterpart: the metric tensor the reinitialisation stages
 * read, gten[c * n + l] for node l (n = the block's node count) and
 * c = 0..5: G_xx, G_xe, G_xz, G_ee, G_ez, G_zz of G_ab = grad xi_a . grad xi_b
 * (xi_a = xi, eta, zeta; grad xi = csi, grad eta = eta, grad zeta = zet).
 * Computed once per grid and passed as LsmBlock.gten it saves three of the
 * nine metric reads per node and stage; without it every reinitialisation
 * call computes it into its workspace. */
int lsm_metric_tensor(const LsmBlock *blk, double *gten, void *stream);

/* Adjoint of lsm_reinit_sussman_eno2 (reverse mode of
 * reinitialisation_test2.F90:42-210 with calc_NormGrad_d.F90:39-2446 and
 * global_lsm.F90:120-133,200-219; the reference has no adjoint): adj_phi0 =
 * (d phi_out / d phi0)^T adj_out.  The iterations are recomputed from phi0
 * (phi_out, optional, receives their result).  The ENO2 / one-sided branches,
 * the upwind side chosen by phi0, the update mask, the lambda switch and the
 * iteration count are taken as in the forward; s(phi0) is differentiated.
 * Where lambda is applied the corrected right-hand side is identically zero,
 * and so is its derivative.  rn_itermax <= 64.  Deterministic.  Workspace:
 * lsm_adjoint_workspace_bytes(blk, LSM_OP_REINIT_ENO2_ADJOINT, rn_itermax). */
int lsm_reinit_sussman_eno2_adjoint(const LsmBlock *blk, const LsmReinitParams *p, const double *phi0,
                                    const double *adj_out, double *adj_phi0, double *phi_out,
                                    void *work, size_t work_bytes, void *stream);

/* calc_sign_ini.F90:1-286 (calc_sign_ini): s(phi0) and the Godunov |grad phi0|. */
int lsm_sign_ini(const LsmBlock *blk, const double *phi0, double epslsm, int m_sphi0,
                 double *sign_out, double *normgrad_out, void *stream);

/* calc_RH_RN.F90:1-244 (calc_RH_RN): rhs = s * (|grad d|_Godunov - 1). */
int lsm_rh_rn(const LsmBlock *blk, const double *d, const double *sign, double *rhs,
              void *stream);

/* calc_lambda_RN.F90:1-122 (calc_lambda_RN, sussman_correction_method = 0). */
int lsm_lambda_rn(const LsmBlock *blk, const double *phi0, const double *normgrad0,
                  const double *d, double timestep, double epslsm, double *lambda,
                  void *stream);

/* calc_lambda_RN.F90:65-120 (calc_lambda_RN, sussman_correction_method = 1):
 * the numerator deltaf(phi0) (d - phi0) / timestep and the denominator
 * deltaf(phi0)^2 |grad phi0| of every node filtered by testfilter.F90:1-124
 * (27-node average, weights 64 / 16 / 4 / 1 divided by aj, at the nodes off the
 * block's faces), lambda = -num / den, 0 where |den| <= 1E-10.  Workspace:
 * lsm_workspace_bytes(blk, LSM_OP_LAMBDA_RN_SIMPSON). */
int lsm_lambda_rn_simpson(const LsmBlock *blk, const double *phi0, const double *normgrad0,
                          const double *d, double timestep, double epslsm, double *lambda,
                          void *work, size_t work_bytes, void *stream);

/* reinitialization_benchmark.F90:1-228 (reinitialization_benchmark, M_Grad = 0,
 * no obstacles, non-hybrid): Godunov pseudo-time reinitialisation; the lambda
 * correction by p->lambda_method (calc_lambda_RN's sussman_correction_method). */
int lsm_reinit_godunov(const LsmBlock *blk, const LsmReinitParams *p,
                       const double *phi0, double *phi_out,
                       void *work, size_t work_bytes, void *stream);

/* TotalWaterVolume.F90:1-297 (TotalWaterVolume): vol_out[0..2] (device) =
 * (K-tetrahedra, second tetrahedra, bulk cells) water volumes; adv_mask
 * (optional) receives the AdvectionNodes interface mask. */
int lsm_water_volume(const LsmBlock *blk, const double *phi, double *vol_out,
                     int32_t *adv_mask, void *work, size_t work_bytes, void *stream);

/* No counterpart in the reference: lsm_water_volume for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same results as lsm_water_volume; no workspace
 * argument and no allocation of any kind: its scratch is a fixed 32 MiB device
 * array of the library (one per GPU, held until the process ends).  A block
 * whose scratch (24 + 4 bytes per 64 x 4 x 8-cell tile, plus one byte per node
 * when adv_mask is asked for) exceeds it is rejected with LSM_ERR_WORKSPACE
 * before any device work.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (several streams or host threads at once) race. */
int lsm_water_volume_noalloc(const LsmBlock *blk, const double *phi, double *vol_out,
                             int32_t *adv_mask, void *stream);

/* levelsetmethod.F90:1-372 (levelsetmethod, non-hybrid Sussman path, no
 * global mass correction): pre-advection water volume (vol_out, device, 3
 * doubles, may be NULL), contravariant velocity from q, numiter_lsm RK3 steps
 * from phi_n, then every rn_freq-th step (ntime % rn_freq == 0) the Sussman
 * reinitialisation and bcond_lsm.  phi is overwritten with phi^(n+1). */
int lsm_levelset_step(const LsmBlock *blk, const LsmReinitParams *p,
                      double *phi, const double *phi_n,
                      const double *u, const double *v, const double *w, int64_t vel_stride,
                      int numiter_lsm, double delti_lsm, int call_levelset,
                      int call_reinit, int rn_freq, int ntime, int mode,
                      double *vol_out, void *work, size_t work_bytes, void *stream);

/* Adjoint of one full level-set step as levelsetmethod.F90:221-326 runs it
 * with numiter_lsm = 1 and a reinitialisation every step: phi^(n+1) =
 * bcond_lsm(reinitialisation_test2(ad_RKTVD3(phi^n))) (the reference has no
 * adjoint).  Writes adj_n = dJ/dphi^n and ADDS dJ/d(U/J) of the step to
 * adj_ucn1..3 when given; phi_np1 (optional) receives the forward result.
 * Composes lsm_rk3_step_adjoint, lsm_reinit_sussman_eno2_adjoint and
 * lsm_bcond_adjoint.  Workspace: lsm_adjoint_workspace_bytes(blk,
 * LSM_OP_STEP_ADJOINT, rn_itermax). */
int lsm_levelset_step_adjoint(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                              const double *ucn1, const double *ucn2, const double *ucn3,
                              double delti_lsm, int mode, const double *adj_np1, double *adj_n,
                              double *adj_ucn1, double *adj_ucn2, double *adj_ucn3, double *phi_np1,
                              void *work, size_t work_bytes, void *stream);

/* ---- batched multi-block form -------------------------------------------
 * Many blocks advanced together.  A connected batch is the block
 * decomposition of one domain (init.F90:489-1068 held on one device): a
 * lattice of blocks whose face sizes match along each lattice slab, `ghost`
 * layers on every face of every block, a face physical exactly where the
 * lattice ends.  An independent batch holds unrelated instances (no ghost
 * layers, six physical faces each).  Every batch array (phi, ucn, metrics,
 * aj) is one device allocation of `total` doubles holding each block's array
 * (extents ni+2*ghost x nj+2*ghost x nk+2*ghost, i fastest, owned nodes at
 * local indices ghost .. ghost+n-1) from its `offset`. */
#define LSM_BATCH_QUEUE     0   /* stage tiles pulled from per-XCD work queues by a persistent grid */
#define LSM_BATCH_GRID      1   /* stage tiles of all blocks, one workgroup each, one launch */
#define LSM_BATCH_PERBLOCK  2   /* the single-block launches, block by block */

typedef struct LsmBatchBlock {
    int32_t ni, nj, nk;               /* owned nodes */
    int32_t lattice[3];               /* position in the block lattice (connected batches) */
    int32_t phys_faces;               /* LSM_FACE_* bits */
    int32_t btype[6];                 /* btype_lsm(1:6) of the block's zone */
    double big_phi;
    double dc, de, dz;
    int64_t offset;                   /* element offset of the block's array in the batch arrays */
    int32_t pitch;                    /* doubles per row of the block's array, >= ni + 2 ghost
                                         (0: ni + 2 ghost); planes hold nj + 2 ghost rows */
} LsmBatchBlock;

typedef struct LsmBatchDesc {
    int32_t nblocks;
    int32_t ghost;                    /* ghost layers (0..3; >= 2 when connected) */
    int32_t connected;                /* 1: one domain (halo fills), 0: independent instances */
    int32_t lattice[3];               /* lattice extents (connected) */
    const LsmBatchBlock *blocks;      /* host array of nblocks entries */
    const double *aj;                 /* device batch array of aj */
    const double *csi[3], *eta[3], *zet[3];  /* device batch metric arrays (NULL: advection from
                                                given U/J only, no reinitialisation) */
    int64_t total;                    /* doubles in every batch array */
} LsmBatchDesc;

typedef struct LsmBatch LsmBatch;

/* Validates the layout and builds the device tables (tile lists, halo and
 * shell enumerations); the batch keeps the aj pointer. */
int lsm_batch_create(const LsmBatchDesc *desc, LsmBatch **out);
int lsm_batch_destroy(LsmBatch *batch);
/* Launch form (LSM_BATCH_*), tile columns (64: a wavefront per row, or 32:
 * two half rows per wavefront; the per-block form takes 64), tile rows (4 or
 * 8), planes per tile and persistent workgroups per CU (0: as many as fit) of
 * the stage kernel.  cfg = {form, cols, rows, kchunk, wg_per_cu}. */
int lsm_batch_set_config(LsmBatch *batch, int form, int tile_cols, int tile_rows, int kchunk, int wg_per_cu);
int lsm_batch_get_config(const LsmBatch *batch, int32_t cfg[5]);
/* Tile columns of the batched reinitialisation stage kernel: 64, 32 or 16
 * (tiles of 64 x 4, 32 x 8 or 16 x 16 nodes), 0 = the default (64).
 * lsm_batch_get_reinit_tile returns the width in use. */
int lsm_batch_set_reinit_tile(LsmBatch *batch, int tile_cols);
int lsm_batch_get_reinit_tile(const LsmBatch *batch);
size_t lsm_batch_workspace_bytes(const LsmBatch *batch, int op);
/* rhs_exchng3_3d.F90:6-359 for every block of a connected batch: each ghost
 * node inside the domain takes the value of the block owning that node. */
int lsm_batch_halo_fill(LsmBatch *batch, double *field, void *stream);
/* rhs_contra_j.F90:2-50 over every node of the batch arrays (the batch's metrics). */
int lsm_batch_contra_velocity(LsmBatch *batch, const double *u, const double *v, const double *w,
                              double *ucn1, double *ucn2, double *ucn3, void *stream);
/* ad_RKTVD3.F90:1-88 for every block: three stages, each followed by
 * bcond_lsm on the physical faces and the halo fill (the reference's
 * rhs_exchng3_3d, ad_RKTVD3.F90:83-84).  phi_n's ghost layers must be
 * filled.  Workspace: lsm_batch_workspace_bytes(batch, LSM_OP_RK3). */
int lsm_batch_rk3_step(LsmBatch *batch, double *phi_np1, const double *phi_n, const double *ucn1,
                       const double *ucn2, const double *ucn3, double delti_lsm, int mode,
                       void *work, size_t work_bytes, void *stream);

/* reinitialisation_test2.F90:1-218 on every block as one rank of the
 * reference runs it: owned nodes updated, ghost layers kept at phi0's values
 * (no exchange inside the loop), |grad d| of ghost nodes 0 in the exit test
 * (a block with ghost layers runs all rn_itermax iterations), one exit test
 * per block.  phi_out's ghost layers receive phi0's.  iters (device, one
 * int32 per block, may be NULL) receives the iterations each block ran.
 * Workspace: lsm_batch_workspace_bytes(batch, LSM_OP_REINIT_ENO2). */
int lsm_batch_reinit_sussman_eno2(LsmBatch *batch, const LsmReinitParams *p, const double *phi0,
                                  double *phi_out, int32_t *iters, void *work, size_t work_bytes,
                                  void *stream);
/* bcond_lsm.F90:2-568 on the physical faces of every block, in place. */
int lsm_batch_bcond(LsmBatch *batch, double *field, void *stream);
/* levelsetmethod.F90:221-326 for every block (numiter_lsm = 1): ad_RKTVD3
 * with bcond_lsm and the halo fill after each stage, then (p != NULL)
 * reinitialisation_test2, bcond_lsm and the halo fill.  Workspace:
 * lsm_batch_workspace_bytes(batch, p ? LSM_OP_LEVELSET_STEP : LSM_OP_RK3). */
int lsm_batch_levelset_step(LsmBatch *batch, const LsmReinitParams *p, double *phi_np1,
                            const double *phi_n, const double *ucn1, const double *ucn2,
                            const double *ucn3, double delti_lsm, int mode, void *work,
                            size_t work_bytes, void *stream);

/* Autotuner (no counterpart in the reference): times the stage-kernel
 * launch configurations (queue form: tile columns 64 / 32, rows 4 / 8, 4 / 8 /
 * 16 planes per tile, persistent workgroups per CU as many as fit or 4; grid
 * form with the same tiles) over `reps` lsm_batch_rk3_step calls on the caller's fields,
 * keeps the fastest and remembers it per shape signature (block extents,
 * ghost layers, connectivity, arithmetic): a later call for a batch of the
 * same shapes applies it without timing unless `force`.  best_ms (optional)
 * receives the winner's milliseconds per step (-1 when taken from the
 * cache).  Synchronises the stream.  phi_np1 is overwritten. */
int lsm_batch_autotune(LsmBatch *batch, double *phi_np1, const double *phi_n, const double *ucn1,
                       const double *ucn2, const double *ucn3, double delti_lsm, int mode, void *work,
                       size_t work_bytes, int reps, int force, double *best_ms, void *stream);
/* Records one lsm_batch_levelset_step (with the current configuration and
 * these arguments, which the graph keeps; p = NULL: advection only) into a
 * HIP graph owned by the batch; lsm_batch_graph_launch replays it on any
 * stream.  Needs a workspace and launch timing switched off.  Any later
 * reconfiguration (lsm_batch_set_config with another configuration,
 * lsm_batch_set_reinit_tile, lsm_batch_autotune) drops the graph: the
 * launch then fails with LSM_ERR_ARG until the step is captured again. */
int lsm_batch_graph_capture(LsmBatch *batch, const LsmReinitParams *p, double *phi_np1,
                            const double *phi_n, const double *ucn1, const double *ucn2,
                            const double *ucn3, double delti_lsm, int mode, void *work,
                            size_t work_bytes);
int lsm_batch_graph_launch(LsmBatch *batch, void *stream);

/* The same step split in two for checkpointed trajectories (no counterpart
 * in the reference): lsm_levelset_step_record runs the forward step from
 * phi_n (phi_np1, optional, receives the result) and keeps in `trace` what
 * the reverse sweep needs (the three RK stage results, the reinitialisation
 * trace, the field after bcond_lsm; lsm_adjoint_workspace_bytes(blk,
 * LSM_OP_STEP_TRACE, rn_itermax) bytes); lsm_levelset_step_adjoint_recorded
 * runs the reverse sweep from such a trace (same blk, p, phi_n, ucn, dt and
 * mode), without recomputing the forward.  Workspace:
 * lsm_adjoint_workspace_bytes(blk, LSM_OP_STEP_ADJOINT_RECORDED, rn_itermax). */
int lsm_levelset_step_record(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                             const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm,
                             int mode, double *phi_np1, void *trace, size_t trace_bytes, void *stream);
int lsm_levelset_step_adjoint_recorded(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                                       const double *ucn1, const double *ucn2, const double *ucn3,
                                       double delti_lsm, int mode, const void *trace, size_t trace_bytes,
                                       const double *adj_np1, double *adj_n, double *adj_ucn1,
                                       double *adj_ucn2, double *adj_ucn3, void *work, size_t work_bytes,
                                       void *stream);

/* ---- multi-instance execution -------------------------------------------
 * No reference counterpart: the reference advances one block per MPI rank.
 * Between lsm_multi_begin and lsm_multi_end (on the calling thread) the
 * entry points lsm_contra_velocity, lsm_rk3_step, lsm_bcond,
 * lsm_rk3_step_adjoint, lsm_bcond_adjoint, lsm_reinit_sussman_eno2(_adjoint),
 * lsm_levelset_step (without volume or hybrid reinitialisation),
 * lsm_levelset_step_record, lsm_levelset_step_adjoint(_recorded) launch
 * nothing: they validate their arguments and record their kernel launches
 * for the current instance (lsm_multi_instance starts the next one; their
 * `stream` arguments are ignored).  Every other entry point fails while
 * recording, and recorded calls need explicit workspaces (one per instance:
 * the instances run at the same time).  lsm_multi_end(stream) requires all
 * instances to have recorded the same kernel sequence (same options, any
 * block shapes) and issues it as one launch per kernel over all instances,
 * in order, on `stream`; lsm_multi_cancel drops a recording.  The results
 * are those of the calls made one after another. */
int lsm_multi_begin(void);
int lsm_multi_instance(void);
int lsm_multi_end(void *stream);
int lsm_multi_cancel(void);

/* ---- host-memory adapter (for Fortran ISO_C_BINDING shims) -------------
 * Same operations on HOST arrays in the reference's layout: the adapter keeps
 * device mirrors of one block and copies arguments in and results out; every
 * call is synchronous.  Metric vectors are the reference's interleaved
 * csi(1:3,i,j,k) arrays and q is q(1:4,i,j,k) with u, v, w in entries 2..4.
 * lsm_host_setup uploads the grid and must precede the other calls; it uploads
 * again only when the extents, the array addresses or a sampled fingerprint of
 * the arrays' contents (about a thousand entries of each) change, so grid arrays
 * rewritten in place are picked up; after editing single entries in place call
 * lsm_host_invalidate, which makes the next lsm_host_setup upload the grid
 * again.  ext = il, iu, jl, ju, kl, ku, igp, jgp, kgp.  The contravariant velocity stays on the device
 * between lsm_host_contra_velocity and the advection calls (one
 * levelsetmethod call computes it once, levelsetmethod.F90:227). */
int lsm_host_setup(const int32_t ext[9], int32_t phys_faces, const int32_t btype[6], double dc,
                   double de, double dz, double big_phi, const double *csi, const double *eta,
                   const double *zet, const double *aj, const double *x, const double *y,
                   const double *z);
/* The blanking regions of the bound block (nblk, li_blk_* and blktype of
 * global_app.F90:30-43): nblk rows of ten values, li_blk_ia, li_blk_ib,
 * li_blk_ja, li_blk_jb, li_blk_ka, li_blk_kb (array indices) and
 * blktype(1:4).  lsm_host_setup clears them; call this after it. */
int lsm_host_blanking(int32_t nblk, const int32_t *boxes);
/* The blanking regions the consumer calls below (lsm_host_phi_gradient_eno2,
 * lsm_host_h_gradient_eno2, lsm_host_calc_h) apply, in the same rows; they
 * stay in force until the next call (nblk = 0: none) and are checked against
 * each call's extents. */
int lsm_host_consumer_blanking(int32_t nblk, const int32_t *boxes);
int lsm_host_invalidate(void);
int lsm_host_release(void);
/* rhs_contra_j.F90:2 */
int lsm_host_contra_velocity(const double *q, double *ucn_aos);
/* calc_RH_AD_WENO3.F90:1 */
int lsm_host_weno3_rhs(const double *phi, double *rhs, const int32_t *adv_mask, int mode);
/* calc_RH_AD (orderAD = 0) on host arrays, as lsm_upwind_rhs */
int lsm_host_upwind_rhs(const double *phi, double *rhs, const int32_t *adv_mask);
/* ad_RKTVD3.F90:1 */
int lsm_host_rk3(double *phi_np1, const double *phi_n, const int32_t *adv_mask, double delti_lsm,
                 int mode);
/* levelsetmethod.F90:1 (the whole call, non-hybrid or hybrid as p->hybrid
 * says, GlobalMassCorrection = .false.): phi, phi_n and q(1:4, ...) are
 * uploaded once, volume, velocity, numiter_lsm RK3 steps, the
 * reinitialisation when mod(ntime, rn_freq) == 0 and bcond_lsm run on the
 * device, and phi^(n+1) and the pre-advection volumes vol[0..2] (K-tetrahedra,
 * second tetrahedra, bulk cells; may be NULL) come back: two field transfers
 * per call instead of two per callee. */
int lsm_host_levelset_step(double *phi, const double *phi_n, const double *q, const LsmReinitParams *p,
                           int numiter_lsm, double delti_lsm, int call_levelset, int call_reinit,
                           int rn_freq, int ntime, int mode, double vol[3]);
/* bcond_lsm.F90:2 */
int lsm_host_bcond(double *phi);
/* calc_NormGrad_d.F90:1 */
int lsm_host_normgrad(const double *d, const double *phizero, int order_bc, int eno_bc,
                      double *normgrad);
/* reinitialisation_test2.F90:1 */
int lsm_host_reinit_eno2(const double *phi0, double *phi_out, const LsmReinitParams *p,
                         int32_t *iters_done);
/* geometric_reinitialisation.F90:1 (GlobalMassCorrection = .false.,
 * TriangulationCase 1), as lsm_reinit_geometric; InterfaceNodesID may be NULL */
int lsm_host_reinit_geometric(const double *phizero, double *phicorrected, int32_t *InterfaceNodesID);
/* calc_sign_ini.F90:1 */
int lsm_host_sign_ini(const double *phi0, double epslsm, int m_sphi0, double *sign_out,
                      double *normgrad_out);
/* calc_RH_RN.F90:1 */
int lsm_host_rh_rn(const double *d, const double *sign, double *rhs);
/* calc_lambda_RN.F90:1 */
int lsm_host_lambda_rn(const double *phi0, const double *normgrad0, const double *d,
                       double timestep, double epslsm, double *lambda);
/* calc_lambda_RN.F90:1 with sussman_correction_method = 1 */
int lsm_host_lambda_rn_simpson(const double *phi0, const double *normgrad0, const double *d,
                               double timestep, double epslsm, double *lambda);
/* reinitialization_benchmark.F90:1 */
int lsm_host_reinit_godunov(const double *phi0, double *phi_out, const LsmReinitParams *p);
/* TotalWaterVolume.F90:1 */
int lsm_host_water_volume(const double *phi, double vol[3], int32_t *adv_mask);
/* The consumers of phi are external subroutines that receive their metrics as
 * arguments (calc_phi_gradient_ENO2.F90:1-11, calc_h.F90:1-8,
 * calc_h_gradient_ENO2.F90:1-11): ext = (il, iu, jl, ju, kl, ku, igp, jgp,
 * kgp), metrics interleaved csi(1:3, ...), outputs interleaved grad(1:nc, ...)
 * and updated in place (vertices keep their values). */
int lsm_host_phi_gradient_eno2(const int32_t ext[9], int32_t phys_faces, double dc, double de, double dz,
                               const double *csi, const double *eta, const double *zet, const double *phi,
                               int order_bc, int eno_bc, double *grad);
int lsm_host_h_gradient_eno2(const int32_t ext[9], int32_t phys_faces, double dc, double de, double dz,
                             const double *csi, const double *eta, const double *zet, const double *h,
                             int order_bc, int eno_bc, double *grad);
int lsm_host_calc_h(const int32_t ext[9], int32_t phys_faces, const double *z, const double *phi,
                    double *h, int32_t *missing);

/* ---- instrumentation (no counterpart in the reference) ------------------
 * Per-kernel launch timing with HIP events recorded on the launch stream,
 * for benchmarks that need a kernel's own duration without a profiler.
 * Kernel ids: */
#define LSM_KT_ADVECT_STAGE   0   /* fused WENO3 + RK stage (interior nodes) */
#define LSM_KT_ADVECT_SHELL   1   /* RK stage on the boundary shell + bcond_lsm */
#define LSM_KT_REINIT_STAGE   2   /* fused ENO2 |grad d| + reinitialisation stage */
#define LSM_KT_VOLUME         3   /* water volume cell kernel */
#define LSM_KT_ADJOINT_STAGE  4   /* fused adjoint of one RK stage (partials + gather) */
#define LSM_KT_BATCH_STAGE    5   /* batched stage tiles of every block (one launch) */
#define LSM_KT_BATCH_REINIT   6   /* batched reinitialisation stage of every block */
#define LSM_KT_REINIT_ADJOINT 7   /* reverse sweep of one reinitialisation stage (tape gather) */
#define LSM_KT_GEOMETRIC      8   /* distances of the geometric reinitialisation */
#define LSM_KT_COUNT          9

/* Turns event recording on (1) or off (0); clears accumulated timings. */
int lsm_timing_enable(int on);

/* Total milliseconds and number of timed launches of kernel `id` since the
 * last enable; waits for the recorded events. */
int lsm_timing_get(int id, double *ms_total, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* LSM_HIP_H */
