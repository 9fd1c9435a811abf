// extern "C" entry points of liblsm_hip.so (declared in include/lsm_hip.h).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "lsm_internal.hpp"
#include "lsm_multi.hpp"

namespace lsm {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return LSM_OK;
    return fail(LSM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int make_blk(const LsmBlock *b, Blk *o, bool need_metrics, bool need_coords, BlankPolicy blanking)
{
    if (recording() && !recordable_here())
        return fail(LSM_ERR_UNSUPPORTED, "this entry point cannot be recorded (lsm_multi_begin .. lsm_multi_end)");
    if (!b) return fail(LSM_ERR_ARG, "LsmBlock is NULL");
    const int ni = b->iu - b->il + 1, nj = b->ju - b->jl + 1, nk = b->ku - b->kl + 1;
    if (ni < 4 || nj < 4 || nk < 4)
        return fail(LSM_ERR_ARG, "block extents %dx%dx%d: every extent must be >= 4", ni, nj, nk);
    if (b->igp != 0 || b->jgp != 0 || b->kgp != 0 || (b->phys_faces & LSM_FACE_ALL) != LSM_FACE_ALL)
        return fail(LSM_ERR_UNSUPPORTED,
                    "only single blocks without ghost layers and with six physical faces are "
                    "supported (igp=%d jgp=%d kgp=%d faces=0x%x)", b->igp, b->jgp, b->kgp, b->phys_faces);
    for (int q = 0; q < 6; ++q)
        if (b->btype[q] < 0 || b->btype[q] > 2)
            return fail(LSM_ERR_ARG, "btype[%d] = %d (expected 0, 1 or 2)", q, b->btype[q]);
    if ((long)ni * nj * nk > (1L << 40)) return fail(LSM_ERR_ARG, "block too large");
    o->e.ni = ni; o->e.nj = nj; o->e.nk = nk;
    o->e.sj = ni; o->e.sk = (long)ni * nj;
    o->e.faces = b->phys_faces & LSM_FACE_ALL;
    for (int q = 0; q < 6; ++q) o->btype[q] = b->btype[q];
    o->dc = b->dc; o->de = b->de; o->dz = b->dz; o->big_phi = b->big_phi;
    o->csi0 = b->csi[0]; o->csi1 = b->csi[1]; o->csi2 = b->csi[2];
    o->eta0 = b->eta[0]; o->eta1 = b->eta[1]; o->eta2 = b->eta[2];
    o->zet0 = b->zet[0]; o->zet1 = b->zet[1]; o->zet2 = b->zet[2];
    o->aj = b->aj; o->x = b->x; o->y = b->y; o->z = b->z;
    o->n = (long)ni * nj * nk;
    o->gten = b->gten;
    if (need_metrics) {
        for (int q = 0; q < 3; ++q)
            if (!b->csi[q] || !b->eta[q] || !b->zet[q])
                return fail(LSM_ERR_ARG, "metric pointer csi/eta/zet[%d] is NULL", q);
        if (!b->aj) return fail(LSM_ERR_ARG, "aj is NULL");
    }
    if (need_coords && (!b->x || !b->y || !b->z)) return fail(LSM_ERR_ARG, "x/y/z is NULL");
    o->bl.n = 0;
    if (b->nblk < 0 || b->nblk > LSM_MAX_BLANK)
        return fail(LSM_ERR_ARG, "nblk = %d (0..%d blanking regions)", b->nblk, LSM_MAX_BLANK);
    if (b->nblk == 0 || blanking == kBlankIgnore) return LSM_OK;
    if (blanking == kBlankReject)
        return fail(LSM_ERR_UNSUPPORTED, "blanking regions (nblk = %d): not implemented by this entry point",
                    b->nblk);
    // boxes in 0-based node indices; the extrapolations of bcond_lsm write
    // the node behind each box face, which must lie in the array
    const int lo[3] = {b->il, b->jl, b->kl}, ext[3] = {ni, nj, nk};
    for (int q = 0; q < b->nblk; ++q) {
        const LsmBlank &B = b->blank[q];
        const int box[6] = {B.ia, B.ib, B.ja, B.jb, B.ka, B.kb};
        for (int d = 0; d < 3; ++d) {
            const int a0 = box[2 * d] - lo[d], a1 = box[2 * d + 1] - lo[d];
            if (a0 < 0 || a0 > a1 || a1 > ext[d] - 1)
                return fail(LSM_ERR_ARG, "blanking region %d: %c range %d..%d outside the block", q, "ijk"[d],
                            box[2 * d], box[2 * d + 1]);
            if (d < 2 && (a0 > ext[d] - 2 || a1 < 1))
                return fail(LSM_ERR_ARG, "blanking region %d: %c range %d..%d needs a node behind each face", q,
                            "ijk"[d], box[2 * d], box[2 * d + 1]);
            o->bl.b[q][2 * d] = a0;
            o->bl.b[q][2 * d + 1] = a1;
        }
        for (int f = 0; f < 4; ++f) o->bl.b[q][6 + f] = B.blktype[f];
    }
    o->bl.n = b->nblk;
    return LSM_OK;
}

// Library-owned device buffers, one per (device, stream, purpose), grown on
// demand and kept until exit.  Keying by stream lets calls on different
// streams run concurrently without sharing scratch; calls on one stream are
// ordered by the stream itself (several host threads issuing calls on the
// same stream at once must serialise them).
static std::mutex g_cache_mu;
static std::map<std::tuple<int, uintptr_t, int>, std::pair<void *, size_t>> g_cache;

int cached_buffer(int purpose, size_t bytes, hipStream_t s, void **ptr)
{
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto &e = g_cache[std::make_tuple(dev, (uintptr_t)s, purpose)];
    if (e.second < bytes) {
        if (e.first) {
            // the buffer may still be in use by queued work of earlier calls on this stream
            rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
            if (rc) return rc;
            (void)hipFree(e.first);
            e.first = nullptr;
            e.second = 0;
        }
        const size_t want = purpose == kBufStore ? bytes + bytes / 4 + 4096 : bytes;
        rc = check_hip(hipMalloc(&e.first, want), purpose == kBufStore ? "hipMalloc(isosurface store)"
                                                                        : "hipMalloc(workspace)");
        if (rc) return rc;
        e.second = want;
    }
    *ptr = e.first;
    return LSM_OK;
}

int default_work(size_t bytes, hipStream_t s, void **ptr)
{
    return cached_buffer(kBufWork, bytes, s, ptr);
}

static int workspace(void *work, size_t work_bytes, size_t need, Carver *c, hipStream_t s)
{
    if (need == 0) need = 256;
    if (work) {
        if (work_bytes < need)
            return fail(LSM_ERR_WORKSPACE, "workspace of %zu bytes, %zu needed", work_bytes, need);
        *c = Carver{(char *)work, work_bytes, 0};
        return LSM_OK;
    }
    // recorded instances run together: one library buffer per stream would be shared by all of them
    if (recording()) return fail(LSM_ERR_ARG, "recorded calls (lsm_multi_begin .. lsm_multi_end) need their own workspace");
    void *p = nullptr;
    int rc = default_work(need, s, &p);
    if (rc) return rc;
    *c = Carver{(char *)p, need, 0};
    return LSM_OK;
}

// ---- launch timing ----
static std::mutex g_t_mu;
static bool g_t_on = false;
struct TPair { int id; hipEvent_t a, b; };
static std::vector<TPair> g_t_pending;
static double g_t_ms[LSM_KT_COUNT];
static long long g_t_n[LSM_KT_COUNT];

bool timing_on() { return g_t_on; }

KTimer::KTimer(int kid, hipStream_t st) : id(kid), s(st)
{
    if (!g_t_on || recording()) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
    (void)hipEventRecord(a, s);
}

KTimer::~KTimer()
{
    if (!a) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(g_t_mu);
    g_t_pending.push_back(TPair{id, a, b});
}

static void drain_timings()
{
    for (auto &p : g_t_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_t_ms[p.id] += ms;
            g_t_n[p.id] += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_t_pending.clear();
}

static size_t field_bytes(const Blk &b) { return ((size_t)b.n * sizeof(double) + 255) & ~(size_t)255; }

// ---- one full step recorded for its adjoint ----
// trace: the three RK stage results (the last is the reinitialisation's
// input), the reinitialisation trace and the field after bcond_lsm
struct StepTrace {
    double *F[3], *Z;
    ReinitTape rt;
};

static size_t step_trace_bytes(const Blk &b, int itermax)
{
    return 4 * field_bytes(b) + reinit_trace_work(b, itermax) + 1024;
}

// the reverse sweep: two adjoint fields, then the reinitialisation adjoint's
// own fields or the advection adjoint's (the two run one after the other)
static size_t step_reverse_bytes(const Blk &b, int itermax)
{
    const size_t r = 4 * field_bytes(b) + 4096;
    const size_t k = 2 * field_bytes(b) + 1024;
    return 2 * field_bytes(b) + (r > k ? r : k) + 1024;
    (void)itermax;
}

static int carve_step(const Blk &b, const LsmReinitParams &p, Carver &c, StepTrace &T, void **ctl)
{
    const size_t fb = (size_t)b.n * sizeof(double);
    for (auto &f : T.F) f = (double *)c.take(fb);
    T.Z = (double *)c.take(fb);
    if (!T.F[0] || !T.F[1] || !T.F[2] || !T.Z) return fail(LSM_ERR_WORKSPACE, "step trace too small");
    double *A, *B;
    return carve_reinit_trace(b, p, T.rt, c, ctl, &A, &B);
}

static int step_record(const Blk &b, const LsmReinitParams &p, const double *phi_n, const double *U1,
                       const double *U2, const double *U3, double dt, int mode, Carver c, StepTrace &T,
                       hipStream_t s)
{
    const size_t fb = (size_t)b.n * sizeof(double);
    for (auto &f : T.F) f = (double *)c.take(fb);
    T.Z = (double *)c.take(fb);
    if (!T.F[0] || !T.F[1] || !T.F[2] || !T.Z) return fail(LSM_ERR_WORKSPACE, "step trace too small");
    int rc;
    if ((rc = launch_rk3_stages(b, phi_n, nullptr, U1, U2, U3, dt, mode, T.F[0], T.F[1], T.F[2], s))) return rc;
    if ((rc = launch_reinit_eno2_trace(b, p, T.F[2], T.rt, T.Z, c, s))) return rc;
    return launch_bcond(b, T.Z, s);
}

static int step_reverse(const Blk &b, const LsmReinitParams &p, const double *phi_n, const double *U1,
                        const double *U2, const double *U3, double dt, int mode, const StepTrace &T,
                        const double *adj_np1, double *adj_n, double *aU1, double *aU2, double *aU3, Carver c,
                        hipStream_t s)
{
    const size_t fb = (size_t)b.n * sizeof(double);
    double *L = (double *)c.take(fb), *LX = (double *)c.take(fb);
    if (!L || !LX) return fail(LSM_ERR_WORKSPACE, "step adjoint workspace too small");
    int rc;
    if ((rc = launch_copy(adj_np1, L, b.n, s))) return rc;
    if ((rc = launch_bcond_adjoint(b, T.Z, L, s))) return rc;
    Carver r = c;
    if ((rc = launch_reinit_eno2_adjoint(b, p, T.F[2], L, LX, nullptr, r, s, &T.rt))) return rc;
    const double *st[3] = {T.F[0], T.F[1], T.F[2]};
    Carver k = c;   // the reinitialisation adjoint's fields are free again
    return launch_rk3_adjoint(b, phi_n, nullptr, U1, U2, U3, dt, mode, LX, adj_n, aU1, aU2, aU3, k, s, st, true);
}

static size_t op_bytes(const Blk &b, int op, int itermax = 0)
{
    switch (op) {
    case LSM_OP_RK3: return field_bytes(b) + 256;
    case LSM_OP_REINIT_ENO2: return reinit_eno2_work(b);
    case LSM_OP_REINIT_GODUNOV: return reinit_godunov_work(b);
    case LSM_OP_WATER_VOLUME: return water_volume_work(b);
    case LSM_OP_RK3_ADJOINT: return rk3_adjoint_work(b);
    case LSM_OP_REINIT_ENO2_ADJOINT: return reinit_eno2_adjoint_work(b, itermax);
    case LSM_OP_STEP_TRACE: return step_trace_bytes(b, itermax);
    case LSM_OP_STEP_ADJOINT_RECORDED: return step_reverse_bytes(b, itermax);
    case LSM_OP_STEP_ADJOINT: return step_trace_bytes(b, itermax) + step_reverse_bytes(b, itermax);
    case LSM_OP_REINIT_GEOMETRIC: return geometric_work(b);
    case LSM_OP_LAMBDA_RN_SIMPSON: return lambda_simpson_work(b);
    case LSM_OP_LEVELSET_STEP: {
        // ucn, phi copy, rk3 scratch, AdvectionNodes mask
        size_t a = 3 * field_bytes(b) + field_bytes(b) + field_bytes(b) + (b.n * 4 + 256) + 1024;
        size_t r = reinit_eno2_work(b);
        size_t v = water_volume_work(b);
        size_t gw = geometric_work(b);
        size_t m = r > v ? r : v;
        return a + (m > gw ? m : gw);
    }
    }
    return 0;
}

}  // namespace lsm

using namespace lsm;

extern "C" {

const char *lsm_version(void) { return "lsm-mi355x 0.1 (gfx950, fp64)"; }

const char *lsm_last_error(void) { return g_err; }

int lsm_timing_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_t_mu);
    drain_timings();
    for (int q = 0; q < LSM_KT_COUNT; ++q) { g_t_ms[q] = 0.0; g_t_n[q] = 0; }
    g_t_on = on != 0;
    return LSM_OK;
}

int lsm_timing_get(int id, double *ms_total, int64_t *launches)
{
    if (id < 0 || id >= LSM_KT_COUNT) return fail(LSM_ERR_ARG, "timing id %d out of range", id);
    std::lock_guard<std::mutex> lk(g_t_mu);
    drain_timings();
    if (ms_total) *ms_total = g_t_ms[id];
    if (launches) *launches = g_t_n[id];
    return LSM_OK;
}

size_t lsm_workspace_bytes(const LsmBlock *blk, int op)
{
    Recordable rec;
    Blk b;
    if (make_blk(blk, &b, false, false, kBlankIgnore)) return 0;
    return op_bytes(b, op);
}

size_t lsm_adjoint_workspace_bytes(const LsmBlock *blk, int op, int rn_itermax)
{
    Recordable rec;
    Blk b;
    if (make_blk(blk, &b, false, false, kBlankIgnore)) return 0;
    if (rn_itermax < 0 || rn_itermax > kMaxTraceIters) return 0;
    return op_bytes(b, op, rn_itermax);
}

int lsm_contra_velocity(const LsmBlock *blk, const double *u, const double *v, const double *w,
                        int64_t vel_stride, double *ucn1, double *ucn2, double *ucn3, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!u || !v || !w || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL array");
    if (vel_stride < 1) return fail(LSM_ERR_ARG, "vel_stride must be >= 1");
    return launch_contra(b, u, v, w, (long)vel_stride, ucn1, ucn2, ucn3, (hipStream_t)stream);
}

int lsm_weno3_rhs(const LsmBlock *blk, const double *phi, const double *ucn1, const double *ucn2,
                  const double *ucn3, const int32_t *adv_mask, double *rhs, int mode, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankUse);
    if (rc) return rc;
    if (!phi || !ucn1 || !ucn2 || !ucn3 || !rhs) return fail(LSM_ERR_ARG, "NULL array");
    if (mode == LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 (LSM_ADVECT_UPWIND1): this entry point covers the WENO3 right-hand side only");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    return launch_weno3_rhs(b, phi, ucn1, ucn2, ucn3, adv_mask, rhs, mode, (hipStream_t)stream);
}

int lsm_upwind_rhs(const LsmBlock *blk, const double *phi, const double *ucn1, const double *ucn2,
                   const double *ucn3, const int32_t *adv_mask, double *rhs, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankIgnore);
    if (rc) return rc;
    if (!phi || !ucn1 || !ucn2 || !ucn3 || !rhs) return fail(LSM_ERR_ARG, "NULL array");
    if (phi == rhs) return fail(LSM_ERR_ARG, "phi and rhs must not alias");
    return launch_upwind_rhs(b, phi, ucn1, ucn2, ucn3, adv_mask, rhs, (hipStream_t)stream);
}

int lsm_bcond(const LsmBlock *blk, double *phi, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankUse);
    if (rc) return rc;
    if (!phi) return fail(LSM_ERR_ARG, "NULL array");
    return launch_bcond(b, phi, (hipStream_t)stream);
}

int lsm_rk3_step(const LsmBlock *blk, double *phi_np1, const double *phi_n, const int32_t *adv_mask,
                 const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm,
                 int mode, void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankUse);
    if (rc) return rc;
    if (!b.aj) return fail(LSM_ERR_ARG, "aj is NULL");
    if (!phi_np1 || !phi_n || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL array");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT && mode != LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_ARG, "bad mode %d", mode);
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_RK3), &c, (hipStream_t)stream);
    if (rc) return rc;
    double *scratch = (double *)c.take(b.n * sizeof(double));
    if (!scratch) return fail(LSM_ERR_WORKSPACE, "workspace too small");
    return launch_rk3(b, phi_np1, phi_n, adv_mask, ucn1, ucn2, ucn3, delti_lsm, mode, scratch,
                      (hipStream_t)stream);
}

int lsm_advect_fold(const LsmBlock *blk, const double *ucn1, const double *ucn2, const double *ucn3,
                    double delti_lsm, double *v1, double *v2, double *v3, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankIgnore);
    if (rc) return rc;
    if (!b.aj) return fail(LSM_ERR_ARG, "aj is NULL");
    if (!ucn1 || !ucn2 || !ucn3 || !v1 || !v2 || !v3) return fail(LSM_ERR_ARG, "NULL array");
    if (!(delti_lsm > 0.0) || !(b.dc > 0.0) || !(b.de > 0.0) || !(b.dz > 0.0))
        return fail(LSM_ERR_UNSUPPORTED, "folded advection needs delti_lsm > 0 and dc, de, dz > 0");
    hipStream_t s = (hipStream_t)stream;
    void *p = nullptr;
    if ((rc = default_work(256, s, &p))) return rc;
    int *bad = (int *)p;
    if ((rc = check_hip(hipMemsetAsync(bad, 0, sizeof(int), s), "hipMemsetAsync"))) return rc;
    if ((rc = launch_fold(b, ucn1, ucn2, ucn3, delti_lsm, v1, v2, v3, bad, s))) return rc;
    int nbad = 0;
    if ((rc = check_hip(hipMemcpyAsync(&nbad, bad, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync")))
        return rc;
    if ((rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize"))) return rc;
    if (nbad) return fail(LSM_ERR_UNSUPPORTED, "aj <= 0 or infinite at %d nodes: folded advection needs 0 < aj < inf", nbad);
    return LSM_OK;
}

int lsm_rk3_step_folded(const LsmBlock *blk, double *phi_np1, const double *phi_n,
                        const int32_t *adv_mask, const double *v1, const double *v2, const double *v3,
                        void *work, size_t work_bytes, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false);
    if (rc) return rc;
    if (!phi_np1 || !phi_n || !v1 || !v2 || !v3) return fail(LSM_ERR_ARG, "NULL array");
    if (phi_np1 == phi_n) return fail(LSM_ERR_ARG, "phi_np1 and phi_n must not alias");
    if (!b.aj) return fail(LSM_ERR_ARG, "aj is NULL");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_RK3), &c, (hipStream_t)stream);
    if (rc) return rc;
    double *scratch = (double *)c.take(b.n * sizeof(double));
    if (!scratch) return fail(LSM_ERR_WORKSPACE, "workspace too small");
    return launch_rk3(b, phi_np1, phi_n, adv_mask, v1, v2, v3, 1.0, LSM_ADVECT_FAST, scratch,
                      (hipStream_t)stream, true);
}

int lsm_rk3_step_adjoint(const LsmBlock *blk, const double *phi_n, const int32_t *adv_mask,
                         const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm,
                         int mode, const double *adj_np1, double *adj_n, double *adj_ucn1,
                         double *adj_ucn2, double *adj_ucn3, void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, false, false);
    if (rc) return rc;
    if (!b.aj) return fail(LSM_ERR_ARG, "aj is NULL");
    if (!phi_n || !ucn1 || !ucn2 || !ucn3 || !adj_np1 || !adj_n) return fail(LSM_ERR_ARG, "NULL array");
    if (adj_n == adj_np1) return fail(LSM_ERR_ARG, "adj_n and adj_np1 must not alias");
    if ((adj_ucn1 != nullptr) != (adj_ucn2 != nullptr) || (adj_ucn1 != nullptr) != (adj_ucn3 != nullptr))
        return fail(LSM_ERR_ARG, "adj_ucn1..3 must all be given or all be NULL");
    if (mode == LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 (LSM_ADVECT_UPWIND1): this entry point covers the WENO3 right-hand side only");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_RK3_ADJOINT), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_rk3_adjoint(b, phi_n, adv_mask, ucn1, ucn2, ucn3, delti_lsm, mode, adj_np1, adj_n,
                              adj_ucn1, adj_ucn2, adj_ucn3, c, (hipStream_t)stream);
}

int lsm_bcond_adjoint(const LsmBlock *blk, const double *phi, double *adj, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, false, false);
    if (rc) return rc;
    if (!phi || !adj) return fail(LSM_ERR_ARG, "NULL array");
    return launch_bcond_adjoint(b, phi, adj, (hipStream_t)stream);
}

int lsm_normgrad_eno2(const LsmBlock *blk, const double *d, const double *phizero, int order_bc,
                      int eno_bc, double *normgrad, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!d || !phizero || !normgrad) return fail(LSM_ERR_ARG, "NULL array");
    if (order_bc < 1 || order_bc > 3) return fail(LSM_ERR_ARG, "order_bc = %d (1, 2 or 3)", order_bc);
    return launch_normgrad(b, d, phizero, order_bc, eno_bc, normgrad, (hipStream_t)stream);
}

static int gradient_entry(const LsmBlock *blk, const double *f, int nc, int order_bc, int eno_bc,
                          double *g1, double *g2, double *g3, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankUse);
    if (rc) return rc;
    for (int q = 0; q < nc; ++q)
        if (!blk->csi[q] || !blk->eta[q] || !blk->zet[q])
            return fail(LSM_ERR_ARG, "metric pointer csi/eta/zet[%d] is NULL", q);
    if (!f || !g1 || !g2 || (nc == 3 && !g3)) return fail(LSM_ERR_ARG, "NULL array");
    if (order_bc < 1 || order_bc > 3) return fail(LSM_ERR_ARG, "order_bc = %d (1, 2 or 3)", order_bc);
    return launch_gradient_eno2(b, f, nc, order_bc, eno_bc, g1, g2, g3, (hipStream_t)stream);
}

int lsm_reinit_geometric(const LsmBlock *blk, const double *phi0, double *phicorrected, int32_t *ids,
                         void *work, size_t work_bytes, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, true, kBlankUse);
    if (rc) return rc;
    if (!phi0 || !phicorrected) return fail(LSM_ERR_ARG, "NULL array");
    if (phi0 == phicorrected) return fail(LSM_ERR_ARG, "phi0 and phicorrected must not alias");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_REINIT_GEOMETRIC), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_geometric(b, phi0, phicorrected, ids, c, (hipStream_t)stream);
}

int lsm_phi_gradient_eno2(const LsmBlock *blk, const double *phi, int order_bc, int eno_bc, double *g1,
                          double *g2, double *g3, void *stream)
{
    return gradient_entry(blk, phi, 3, order_bc, eno_bc, g1, g2, g3, stream);
}

int lsm_h_gradient_eno2(const LsmBlock *blk, const double *h, int order_bc, int eno_bc, double *g1,
                        double *g2, void *stream)
{
    return gradient_entry(blk, h, 2, order_bc, eno_bc, g1, g2, nullptr, stream);
}

int lsm_calc_h(const LsmBlock *blk, const double *z, const double *phi, double *h, int32_t *missing,
               void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankUse);
    if (rc) return rc;
    if (!z || !phi || !h) return fail(LSM_ERR_ARG, "NULL array");
    return launch_free_surface(b, z, phi, h, nullptr, missing, (hipStream_t)stream);
}

int lsm_free_surface(const LsmBlock *blk, const double *z, const double *phi, double *zfs, int32_t *missing,
                     void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false);
    if (rc) return rc;
    if (!z || !phi || !zfs) return fail(LSM_ERR_ARG, "NULL array");
    return launch_free_surface(b, z, phi, nullptr, zfs, missing, (hipStream_t)stream);
}

int lsm_free_surface_adjoint(const LsmBlock *blk, const double *z, const double *phi, const double *adj_zfs,
                             double *adj_phi, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false);
    if (rc) return rc;
    if (!z || !phi || !adj_zfs || !adj_phi) return fail(LSM_ERR_ARG, "NULL array");
    return launch_free_surface_adjoint(b, z, phi, adj_zfs, adj_phi, (hipStream_t)stream);
}

static int check_reinit(const LsmReinitParams *p)
{
    if (!p) return fail(LSM_ERR_ARG, "LsmReinitParams is NULL");
    if (p->order_bc < 1 || p->order_bc > 3) return fail(LSM_ERR_ARG, "order_bc = %d (1, 2 or 3)", p->order_bc);
    if (p->rn_itermax < 0) return fail(LSM_ERR_ARG, "rn_itermax < 0");
    if (!(p->epslsm > 0.0)) return fail(LSM_ERR_ARG, "epslsm must be > 0");
    if (p->lambda_method != 0 && p->lambda_method != 1)
        return fail(LSM_ERR_ARG, "lambda_method = %d (0 or 1)", p->lambda_method);
    return LSM_OK;
}

int lsm_reinit_sussman_eno2(const LsmBlock *blk, const LsmReinitParams *p, const double *phi0,
                            double *phi_out, int32_t *iters_done, void *work, size_t work_bytes,
                            void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if ((rc = check_reinit(p))) return rc;
    if (!phi0 || !phi_out) return fail(LSM_ERR_ARG, "NULL array");
    if (phi0 == phi_out) return fail(LSM_ERR_ARG, "phi0 and phi_out must not alias");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_REINIT_ENO2), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_reinit_eno2(b, *p, phi0, phi_out, iters_done, c, (hipStream_t)stream);
}

int lsm_metric_tensor(const LsmBlock *blk, double *gten, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!gten) return fail(LSM_ERR_ARG, "gten is NULL");
    return launch_metric_tensor(b, gten, (hipStream_t)stream);
}

int lsm_reinit_sussman_eno2_adjoint(const LsmBlock *blk, const LsmReinitParams *p, const double *phi0,
                                    const double *adj_out, double *adj_phi0, double *phi_out,
                                    void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if ((rc = check_reinit(p))) return rc;
    if (p->rn_itermax > kMaxTraceIters)
        return fail(LSM_ERR_UNSUPPORTED, "rn_itermax = %d: the adjoint keeps at most %d iterations",
                    p->rn_itermax, kMaxTraceIters);
    if (!phi0 || !adj_out || !adj_phi0) return fail(LSM_ERR_ARG, "NULL array");
    if (adj_phi0 == adj_out) return fail(LSM_ERR_ARG, "adj_phi0 and adj_out must not alias");
    if (phi_out && (phi_out == phi0 || phi_out == adj_out || phi_out == adj_phi0))
        return fail(LSM_ERR_ARG, "phi_out must not alias the other arrays");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_REINIT_ENO2_ADJOINT, p->rn_itermax), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_reinit_eno2_adjoint(b, *p, phi0, adj_out, adj_phi0, phi_out, c, (hipStream_t)stream);
}

static int check_step_adjoint(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                              const double *ucn1, const double *ucn2, const double *ucn3, int mode, Blk &b)
{
    int rc = make_blk(blk, &b, true, false);
    if (rc) return rc;
    if ((rc = check_reinit(p))) return rc;
    if (p->hybrid)
        return fail(LSM_ERR_UNSUPPORTED, "hybrid (geometric) reinitialisation: the step adjoint covers the Sussman path only");
    if (p->rn_itermax > kMaxTraceIters)
        return fail(LSM_ERR_UNSUPPORTED, "rn_itermax = %d: the adjoint keeps at most %d iterations",
                    p->rn_itermax, kMaxTraceIters);
    if (!phi_n || !ucn1 || !ucn2 || !ucn3) return fail(LSM_ERR_ARG, "NULL array");
    if (mode == LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_UNSUPPORTED, "orderAD = 0 (LSM_ADVECT_UPWIND1): this entry point covers the WENO3 right-hand side only");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT) return fail(LSM_ERR_ARG, "bad mode %d", mode);
    return LSM_OK;
}

static int check_adj_args(const double *adj_np1, double *adj_n, double *a1, double *a2, double *a3)
{
    if (!adj_np1 || !adj_n) return fail(LSM_ERR_ARG, "NULL array");
    if (adj_n == adj_np1) return fail(LSM_ERR_ARG, "adj_n and adj_np1 must not alias");
    if ((a1 != nullptr) != (a2 != nullptr) || (a1 != nullptr) != (a3 != nullptr))
        return fail(LSM_ERR_ARG, "adj_ucn1..3 must all be given or all be NULL");
    return LSM_OK;
}

int lsm_levelset_step_adjoint(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                              const double *ucn1, const double *ucn2, const double *ucn3,
                              double delti_lsm, int mode, const double *adj_np1, double *adj_n,
                              double *adj_ucn1, double *adj_ucn2, double *adj_ucn3, double *phi_np1,
                              void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = check_step_adjoint(blk, p, phi_n, ucn1, ucn2, ucn3, mode, b);
    if (rc) return rc;
    if ((rc = check_adj_args(adj_np1, adj_n, adj_ucn1, adj_ucn2, adj_ucn3))) return rc;
    hipStream_t s = (hipStream_t)stream;
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_STEP_ADJOINT, p->rn_itermax), &c, (hipStream_t)stream);
    if (rc) return rc;
    // forward, recorded into the first part of the workspace, then the reverse sweep
    StepTrace T;
    Carver tr{c.base, step_trace_bytes(b, p->rn_itermax), 0};
    if ((rc = step_record(b, *p, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, tr, T, s))) return rc;
    if (phi_np1 && (rc = launch_copy(T.Z, phi_np1, b.n, s))) return rc;
    const size_t off = (step_trace_bytes(b, p->rn_itermax) + 255) & ~(size_t)255;
    Carver rv{c.base + off, c.size - off, 0};
    return step_reverse(b, *p, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, T, adj_np1, adj_n, adj_ucn1, adj_ucn2,
                        adj_ucn3, rv, s);
}

int lsm_levelset_step_record(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                             const double *ucn1, const double *ucn2, const double *ucn3, double delti_lsm,
                             int mode, double *phi_np1, void *trace, size_t trace_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = check_step_adjoint(blk, p, phi_n, ucn1, ucn2, ucn3, mode, b);
    if (rc) return rc;
    if (!trace) return fail(LSM_ERR_ARG, "NULL trace");
    if (trace_bytes < step_trace_bytes(b, p->rn_itermax))
        return fail(LSM_ERR_WORKSPACE, "trace of %zu bytes, %zu needed", trace_bytes, step_trace_bytes(b, p->rn_itermax));
    hipStream_t s = (hipStream_t)stream;
    StepTrace T;
    if ((rc = step_record(b, *p, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, Carver{(char *)trace, trace_bytes, 0}, T, s)))
        return rc;
    if (phi_np1 && (rc = launch_copy(T.Z, phi_np1, b.n, s))) return rc;
    return LSM_OK;
}

int lsm_levelset_step_adjoint_recorded(const LsmBlock *blk, const LsmReinitParams *p, const double *phi_n,
                                       const double *ucn1, const double *ucn2, const double *ucn3,
                                       double delti_lsm, int mode, const void *trace, size_t trace_bytes,
                                       const double *adj_np1, double *adj_n, double *adj_ucn1, double *adj_ucn2,
                                       double *adj_ucn3, void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    Blk b;
    int rc = check_step_adjoint(blk, p, phi_n, ucn1, ucn2, ucn3, mode, b);
    if (rc) return rc;
    if ((rc = check_adj_args(adj_np1, adj_n, adj_ucn1, adj_ucn2, adj_ucn3))) return rc;
    if (!trace) return fail(LSM_ERR_ARG, "NULL trace");
    if (trace_bytes < step_trace_bytes(b, p->rn_itermax)) return fail(LSM_ERR_WORKSPACE, "trace too small");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_STEP_ADJOINT_RECORDED, p->rn_itermax), &c, (hipStream_t)stream);
    if (rc) return rc;
    StepTrace T;
    Carver tr{(char *)const_cast<void *>(trace), trace_bytes, 0};
    if ((rc = carve_step(b, *p, tr, T, nullptr))) return rc;
    return step_reverse(b, *p, phi_n, ucn1, ucn2, ucn3, delti_lsm, mode, T, adj_np1, adj_n, adj_ucn1, adj_ucn2,
                        adj_ucn3, c, (hipStream_t)stream);
}

int lsm_sign_ini(const LsmBlock *blk, const double *phi0, double epslsm, int m_sphi0,
                 double *sign_out, double *normgrad_out, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!phi0 || !sign_out || !normgrad_out) return fail(LSM_ERR_ARG, "NULL array");
    return launch_sign_ini(b, phi0, epslsm, m_sphi0, sign_out, normgrad_out, (hipStream_t)stream);
}

int lsm_rh_rn(const LsmBlock *blk, const double *d, const double *sign, double *rhs, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!d || !sign || !rhs) return fail(LSM_ERR_ARG, "NULL array");
    return launch_rh_rn(b, d, sign, rhs, (hipStream_t)stream);
}

int lsm_lambda_rn(const LsmBlock *blk, const double *phi0, const double *normgrad0, const double *d,
                  double timestep, double epslsm, double *lambda, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, false, kBlankIgnore);
    if (rc) return rc;
    if (!phi0 || !normgrad0 || !d || !lambda) return fail(LSM_ERR_ARG, "NULL array");
    return launch_lambda_rn(b, phi0, normgrad0, d, timestep, epslsm, lambda, (hipStream_t)stream);
}

int lsm_lambda_rn_simpson(const LsmBlock *blk, const double *phi0, const double *normgrad0, const double *d,
                          double timestep, double epslsm, double *lambda, void *work, size_t work_bytes,
                          void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if (!phi0 || !normgrad0 || !d || !lambda) return fail(LSM_ERR_ARG, "NULL array");
    if (lambda == phi0 || lambda == normgrad0 || lambda == d)
        return fail(LSM_ERR_ARG, "lambda must not alias an input (the filter reads neighbours)");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_LAMBDA_RN_SIMPSON), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_lambda_rn_simpson(b, phi0, normgrad0, d, timestep, epslsm, lambda, c, (hipStream_t)stream);
}

int lsm_reinit_godunov(const LsmBlock *blk, const LsmReinitParams *p, const double *phi0,
                       double *phi_out, void *work, size_t work_bytes, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, true, false, kBlankIgnore);
    if (rc) return rc;
    if ((rc = check_reinit(p))) return rc;
    if (!phi0 || !phi_out) return fail(LSM_ERR_ARG, "NULL array");
    if (phi0 == phi_out) return fail(LSM_ERR_ARG, "phi0 and phi_out must not alias");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_REINIT_GODUNOV), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_reinit_godunov(b, *p, phi0, phi_out, c, (hipStream_t)stream);
}

int lsm_water_volume(const LsmBlock *blk, const double *phi, double *vol_out, int32_t *adv_mask,
                     void *work, size_t work_bytes, void *stream)
{
    Blk b;
    int rc = make_blk(blk, &b, false, true, kBlankUse);
    if (rc) return rc;
    if (!phi || !vol_out) return fail(LSM_ERR_ARG, "NULL array");
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_WATER_VOLUME), &c, (hipStream_t)stream);
    if (rc) return rc;
    return launch_water_volume(b, phi, vol_out, adv_mask, c, (hipStream_t)stream);
}

int lsm_levelset_step(const LsmBlock *blk, const LsmReinitParams *p, double *phi,
                      const double *phi_n, const double *u, const double *v, const double *w,
                      int64_t vel_stride, int numiter_lsm, double delti_lsm, int call_levelset,
                      int call_reinit, int rn_freq, int ntime, int mode, double *vol_out,
                      void *work, size_t work_bytes, void *stream)
{
    Recordable rec;
    const bool hybrid = p && p->hybrid && call_reinit;
    if (recording() && (vol_out || hybrid))
        return fail(LSM_ERR_UNSUPPORTED, "the water volume and the hybrid reinitialisation cannot be recorded");
    Blk b;
    int rc = make_blk(blk, &b, true, vol_out != nullptr || hybrid, kBlankUse);
    if (rc) return rc;
    if (!phi || !phi_n) return fail(LSM_ERR_ARG, "NULL array");
    if (call_levelset && (!u || !v || !w)) return fail(LSM_ERR_ARG, "NULL velocity");
    if (vel_stride < 1) return fail(LSM_ERR_ARG, "vel_stride must be >= 1");
    if (numiter_lsm < 0) return fail(LSM_ERR_ARG, "numiter_lsm < 0");
    if (mode != LSM_ADVECT_FAST && mode != LSM_ADVECT_EXACT && mode != LSM_ADVECT_UPWIND1)
        return fail(LSM_ERR_ARG, "bad mode %d", mode);
    const bool do_reinit = call_reinit && rn_freq > 0 && ntime % rn_freq == 0;
    if (do_reinit && !hybrid && (rc = check_reinit(p))) return rc;
    hipStream_t s = (hipStream_t)stream;
    Carver c;
    rc = workspace(work, work_bytes, op_bytes(b, LSM_OP_LEVELSET_STEP), &c, (hipStream_t)stream);
    if (rc) return rc;
    const size_t fb = b.n * sizeof(double);
    double *U1 = (double *)c.take(fb), *U2 = (double *)c.take(fb), *U3 = (double *)c.take(fb);
    double *cur = (double *)c.take(fb), *tmp = (double *)c.take(fb);
    int32_t *mask = hybrid ? (int32_t *)c.take(b.n * 4) : nullptr;
    if (!U1 || !U2 || !U3 || !cur || !tmp || (hybrid && !mask))
        return fail(LSM_ERR_WORKSPACE, "workspace too small");
    // levelsetmethod.F90:178-196: pre-advection water volume and the
    // AdvectionNodes band, which restricts the advection on the hybrid path
    // only (levelsetmethod.F90:198-199)
    if (vol_out || hybrid) {
        Carver vc = c;
        if ((rc = launch_water_volume(b, phi, vol_out, mask, vc, s))) return rc;
    }
    // levelsetmethod.F90:221-243: contravariant velocity, numiter_lsm RK3 steps from phi_n
    if (call_levelset) {
        if ((rc = launch_contra(b, u, v, w, (long)vel_stride, U1, U2, U3, s))) return rc;
        if ((rc = launch_copy(phi_n, cur, b.n, s))) return rc;
        for (int it = 0; it < numiter_lsm; ++it) {
            if ((rc = launch_rk3(b, phi, cur, mask, U1, U2, U3, delti_lsm, mode, tmp, s))) return rc;
            if (it + 1 < numiter_lsm && (rc = launch_copy(phi, cur, b.n, s))) return rc;
        }
    }
    // levelsetmethod.F90:290-326: reinitialisation (geometric on the hybrid
    // path, Sussman otherwise), bcond_lsm, phi = phicorrected
    if (do_reinit) {
        Carver rcv = c;
        if (hybrid) {
            // phicorrected: +-9.9 away from the zero set, the advected phi
            // (0 without advection) where |phi| <= eps (levelsetmethod.F90:300-301)
            if ((rc = launch_hybrid_init(phi, call_levelset ? phi : nullptr, tmp, b.n, s))) return rc;
            if ((rc = launch_geometric(b, phi, tmp, nullptr, rcv, s))) return rc;
        } else {
            if ((rc = launch_reinit_eno2(b, *p, phi, tmp, nullptr, rcv, s))) return rc;
        }
        if ((rc = launch_bcond(b, tmp, s))) return rc;
        if ((rc = launch_copy(tmp, phi, b.n, s))) return rc;
    }
    return LSM_OK;
}

}  // extern "C"
