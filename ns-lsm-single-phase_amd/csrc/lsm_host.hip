// Host-memory adapter of the C ABI for callers whose arrays live in host
// memory -- the Fortran ISO_C_BINDING shims that replace the reference's
// level-set routines inside levelsetmethod.  The adapter keeps device mirrors
// of one block (metrics uploaded once per grid, SoA), the contravariant
// velocity of the last rhs_contra_j call, and staging buffers; every call
// copies its array arguments in, runs the device entry point on the default
// stream and copies results out before returning (the reference reads results
// right after each call).  Output arrays are copied in as well, so entries a
// routine does not write (boundary nodes of a right-hand side, say) keep the
// caller's values, as they do in the reference.
#include <string.h>

#include <mutex>
#include <vector>

#include "lsm_internal.hpp"

namespace lsm {
namespace {

struct HostState {
    // identity of the uploaded grid
    int ext[9] = {0};
    const void *key[7] = {nullptr};
    uint64_t sig = 0;                // sampled fingerprint of the grid arrays' contents
    bool ready = false;
    LsmBlock blk{};
    long n = 0;
    double *dm[13] = {nullptr};      // csi0..2, eta0..2, zet0..2, aj, x, y, z
    double *ucn[3] = {nullptr};
    bool have_ucn = false;
    double *f[4] = {nullptr};        // staging fields
    int32_t *mask = nullptr;
    double *vol = nullptr;
    int32_t *iters = nullptr;
    double *q = nullptr;             // q(1:4, ...) of a whole-levelsetmethod call
    void *lswork = nullptr;          // its workspace
    size_t lswork_bytes = 0;
    std::vector<double> aos;         // host transposition buffer
};

std::mutex g_mu;
HostState g_h;

void release()
{
    for (auto &p : g_h.dm) { if (p) (void)hipFree(p); p = nullptr; }
    for (auto &p : g_h.ucn) { if (p) (void)hipFree(p); p = nullptr; }
    for (auto &p : g_h.f) { if (p) (void)hipFree(p); p = nullptr; }
    if (g_h.mask) (void)hipFree(g_h.mask);
    if (g_h.vol) (void)hipFree(g_h.vol);
    if (g_h.iters) (void)hipFree(g_h.iters);
    g_h.mask = nullptr; g_h.vol = nullptr; g_h.iters = nullptr;
    if (g_h.q) (void)hipFree(g_h.q);
    if (g_h.lswork) (void)hipFree(g_h.lswork);
    g_h.q = nullptr; g_h.lswork = nullptr; g_h.lswork_bytes = 0;
    g_h.ready = false;
    g_h.have_ucn = false;
}

int alloc(void **p, size_t bytes) { return check_hip(hipMalloc(p, bytes), "hipMalloc(host adapter)"); }

// FNV-1a over up to 1024 evenly spaced entries (and the last one) of each grid
// array: an in-place rewrite of a grid array (new metrics at the same
// address) changes it; edits of single entries may not, for those the caller
// calls lsm_host_invalidate
uint64_t grid_signature(const double *const arr[7], const long len[7])
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const double v) {
        uint64_t u;
        memcpy(&u, &v, sizeof u);
        for (int q = 0; q < 8; ++q) { h ^= (u >> (8 * q)) & 0xff; h *= 1099511628211ull; }
    };
    for (int a = 0; a < 7; ++a) {
        if (!arr[a] || len[a] <= 0) continue;
        const long stride = len[a] > 1024 ? len[a] / 1024 : 1;
        for (long q = 0; q < len[a]; q += stride) mix(arr[a][q]);
        mix(arr[a][len[a] - 1]);
    }
    return h;
}

int up(double *d, const double *h, long n)
{
    return check_hip(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int down(double *h, const double *d, long n)
{
    return check_hip(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

// interleaved (3, n) host array -> three device arrays
int up_aos(double *d0, double *d1, double *d2, const double *h, long n)
{
    g_h.aos.resize((size_t)n);
    double *dst[3] = {d0, d1, d2};
    for (int c = 0; c < 3; ++c) {
        for (long l = 0; l < n; ++l) g_h.aos[l] = h[3 * l + c];
        int rc = up(dst[c], g_h.aos.data(), n);
        if (rc) return rc;
    }
    return LSM_OK;
}

int sync() { return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

// interleaved (nc, n) host array <-> nc device arrays
int up_aosn(double *const *d, int nc, const double *h, long n, std::vector<double> &buf)
{
    buf.resize((size_t)n);
    for (int c = 0; c < nc; ++c) {
        for (long l = 0; l < n; ++l) buf[l] = h[nc * l + c];
        int rc = up(d[c], buf.data(), n);
        if (rc) return rc;
    }
    return LSM_OK;
}

int down_aosn(double *h, double *const *d, int nc, long n, std::vector<double> &buf)
{
    buf.resize((size_t)n);
    for (int c = 0; c < nc; ++c) {
        int rc = down(buf.data(), d[c], n);
        if (rc) return rc;
        for (long l = 0; l < n; ++l) h[nc * l + c] = buf[l];
    }
    return LSM_OK;
}

// The consumers of phi (calc_phi_gradient_ENO2, calc_h, calc_h_gradient_ENO2)
// are external subroutines with their own argument lists, called outside
// levelsetmethod: they get a mirror of their own, keyed by the extents and
// the metric arrays they receive.
struct ConsumerState {
    int ext[9] = {0};
    int faces = -1;
    const void *key[3] = {nullptr};
    bool have_metrics = false;
    long n = 0;
    double *dm[9] = {nullptr};
    double *f[4] = {nullptr};
    int32_t *cnt = nullptr;
    std::vector<double> buf;
    int32_t nblk = 0;              // blanking regions of the consumer calls (lsm_host_consumer_blanking)
    LsmBlank blank[LSM_MAX_BLANK] = {};
};

ConsumerState g_c;

// rows of ten values (ia, ib, ja, jb, ka, kb, blktype(1:4)) into the block
int set_blanking(LsmBlock &b, int32_t nblk, const int32_t *boxes)
{
    if (nblk < 0 || nblk > LSM_MAX_BLANK) return fail(LSM_ERR_ARG, "nblk = %d (0..%d)", nblk, LSM_MAX_BLANK);
    if (nblk > 0 && !boxes) return fail(LSM_ERR_ARG, "NULL boxes");
    b.nblk = nblk;
    for (int q = 0; q < nblk; ++q) {
        const int32_t *v = boxes + 10 * q;
        LsmBlank &B = b.blank[q];
        B.ia = v[0]; B.ib = v[1]; B.ja = v[2]; B.jb = v[3]; B.ka = v[4]; B.kb = v[5];
        for (int f = 0; f < 4; ++f) B.blktype[f] = v[6 + f];
    }
    return LSM_OK;
}

void consumer_release()
{
    for (auto &p : g_c.dm) { if (p) (void)hipFree(p); p = nullptr; }
    for (auto &p : g_c.f) { if (p) (void)hipFree(p); p = nullptr; }
    if (g_c.cnt) (void)hipFree(g_c.cnt);
    g_c.cnt = nullptr;
    g_c.n = 0;
    g_c.have_metrics = false;
}

// validates the extents, (re)allocates the mirror and, when csi is given,
// uploads the metrics unless the same arrays were uploaded last time
int consumer_block(const int32_t ext[9], int32_t faces, double dc, double de, double dz, const double *csi,
                   const double *eta, const double *zet, LsmBlock *b)
{
    if (!ext) return fail(LSM_ERR_ARG, "NULL argument");
    *b = LsmBlock{};
    b->il = ext[0]; b->iu = ext[1]; b->jl = ext[2]; b->ju = ext[3]; b->kl = ext[4]; b->ku = ext[5];
    b->igp = ext[6]; b->jgp = ext[7]; b->kgp = ext[8];
    b->phys_faces = faces;
    for (int q = 0; q < 6; ++q) b->btype[q] = 1;
    b->dc = dc; b->de = de; b->dz = dz;
    b->nblk = g_c.nblk;
    for (int q = 0; q < LSM_MAX_BLANK; ++q) b->blank[q] = g_c.blank[q];
    Blk chk;
    int rc = make_blk(b, &chk, false, false, kBlankUse);
    if (rc) return rc;
    if (memcmp(g_c.ext, ext, sizeof g_c.ext) != 0 || g_c.faces != faces || !g_c.f[0]) {
        consumer_release();
        for (auto &p : g_c.dm) if ((rc = alloc((void **)&p, chk.n * sizeof(double)))) return rc;
        for (auto &p : g_c.f) if ((rc = alloc((void **)&p, chk.n * sizeof(double)))) return rc;
        if ((rc = alloc((void **)&g_c.cnt, sizeof(int32_t)))) return rc;
        memcpy(g_c.ext, ext, sizeof g_c.ext);
        g_c.faces = faces;
        g_c.n = chk.n;
    }
    if (csi) {
        const void *key[3] = {csi, eta, zet};
        if (!g_c.have_metrics || memcmp(g_c.key, key, sizeof key) != 0) {
            if ((rc = up_aosn(g_c.dm, 3, csi, g_c.n, g_c.buf)) || (rc = up_aosn(g_c.dm + 3, 3, eta, g_c.n, g_c.buf)) ||
                (rc = up_aosn(g_c.dm + 6, 3, zet, g_c.n, g_c.buf)))
                return rc;
            memcpy(g_c.key, key, sizeof key);
            g_c.have_metrics = true;
        }
        for (int c = 0; c < 3; ++c) {
            b->csi[c] = g_c.dm[c];
            b->eta[c] = g_c.dm[3 + c];
            b->zet[c] = g_c.dm[6 + c];
        }
    }
    return LSM_OK;
}

int consumer_gradient(const int32_t ext[9], int32_t faces, double dc, double de, double dz, const double *csi,
                      const double *eta, const double *zet, const double *f, int nc, int order_bc, int eno_bc,
                      double *grad)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!csi || !eta || !zet || !f || !grad) return fail(LSM_ERR_ARG, "NULL argument");
    LsmBlock b;
    int rc = consumer_block(ext, faces, dc, de, dz, csi, eta, zet, &b);
    if (rc) return rc;
    const long n = g_c.n;
    if ((rc = up(g_c.f[0], f, n))) return rc;
    if ((rc = up_aosn(g_c.f + 1, nc, grad, n, g_c.buf))) return rc;
    rc = nc == 3 ? lsm_phi_gradient_eno2(&b, g_c.f[0], order_bc, eno_bc, g_c.f[1], g_c.f[2], g_c.f[3], nullptr)
                 : lsm_h_gradient_eno2(&b, g_c.f[0], order_bc, eno_bc, g_c.f[1], g_c.f[2], nullptr);
    if (rc) return rc;
    if ((rc = sync())) return rc;
    return down_aosn(grad, g_c.f + 1, nc, n, g_c.buf);
}

int need_ready()
{
    if (!g_h.ready) return fail(LSM_ERR_ARG, "lsm_host_setup has not been called");
    return LSM_OK;
}


}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" {

int lsm_host_setup(const int32_t ext[9], int32_t phys_faces, const int32_t btype[6], double dc,
                   double de, double dz, double big_phi, const double *csi, const double *eta,
                   const double *zet, const double *aj, const double *x, const double *y,
                   const double *z)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ext || !btype || !csi || !eta || !zet || !aj) return fail(LSM_ERR_ARG, "NULL argument");
    const void *key[7] = {csi, eta, zet, aj, x, y, z};
    LsmBlock b{};
    b.il = ext[0]; b.iu = ext[1]; b.jl = ext[2]; b.ju = ext[3]; b.kl = ext[4]; b.ku = ext[5];
    b.igp = ext[6]; b.jgp = ext[7]; b.kgp = ext[8];
    b.phys_faces = phys_faces;
    for (int q = 0; q < 6; ++q) b.btype[q] = btype[q];
    b.dc = dc; b.de = de; b.dz = dz; b.big_phi = big_phi;
    Blk chk;
    int rc = make_blk(&b, &chk, false, false);
    if (rc) return rc;
    const double *arr[7] = {csi, eta, zet, aj, x, y, z};
    const long len[7] = {3 * chk.n, 3 * chk.n, 3 * chk.n, chk.n, x ? chk.n : 0, y ? chk.n : 0, z ? chk.n : 0};
    const uint64_t sig = grid_signature(arr, len);
    // the device mirror is reused while the extents, the array addresses and
    // the contents' fingerprint are those of the upload
    const bool same = g_h.ready && memcmp(g_h.ext, ext, sizeof g_h.ext) == 0 &&
                      memcmp(g_h.key, key, sizeof key) == 0 && g_h.sig == sig;
    if (!same) {
        release();
        const long n = chk.n;
        for (auto &p : g_h.dm) if ((rc = alloc((void **)&p, n * sizeof(double)))) return rc;
        for (auto &p : g_h.ucn) if ((rc = alloc((void **)&p, n * sizeof(double)))) return rc;
        for (auto &p : g_h.f) if ((rc = alloc((void **)&p, n * sizeof(double)))) return rc;
        if ((rc = alloc((void **)&g_h.mask, n * sizeof(int32_t)))) return rc;
        if ((rc = alloc((void **)&g_h.vol, 4 * sizeof(double)))) return rc;
        if ((rc = alloc((void **)&g_h.iters, sizeof(int32_t)))) return rc;
        if ((rc = up_aos(g_h.dm[0], g_h.dm[1], g_h.dm[2], csi, n))) return rc;
        if ((rc = up_aos(g_h.dm[3], g_h.dm[4], g_h.dm[5], eta, n))) return rc;
        if ((rc = up_aos(g_h.dm[6], g_h.dm[7], g_h.dm[8], zet, n))) return rc;
        if ((rc = up(g_h.dm[9], aj, n))) return rc;
        if (x && y && z) {
            if ((rc = up(g_h.dm[10], x, n)) || (rc = up(g_h.dm[11], y, n)) || (rc = up(g_h.dm[12], z, n)))
                return rc;
        }
        memcpy(g_h.ext, ext, sizeof g_h.ext);
        memcpy(g_h.key, key, sizeof key);
        g_h.sig = sig;
        g_h.n = n;
        g_h.ready = true;
    }
    for (int c = 0; c < 3; ++c) {
        b.csi[c] = g_h.dm[c];
        b.eta[c] = g_h.dm[3 + c];
        b.zet[c] = g_h.dm[6 + c];
    }
    b.aj = g_h.dm[9];
    b.x = (x && y && z) ? g_h.dm[10] : nullptr;
    b.y = (x && y && z) ? g_h.dm[11] : nullptr;
    b.z = (x && y && z) ? g_h.dm[12] : nullptr;
    g_h.blk = b;
    return LSM_OK;
}

int lsm_host_blanking(int32_t nblk, const int32_t *boxes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (nblk < 0 || nblk > LSM_MAX_BLANK) return fail(LSM_ERR_ARG, "nblk = %d (0..%d)", nblk, LSM_MAX_BLANK);
    if (nblk > 0 && !boxes) return fail(LSM_ERR_ARG, "NULL boxes");
    LsmBlock b = g_h.blk;
    if ((rc = set_blanking(b, nblk, boxes))) return rc;
    Blk chk;
    if ((rc = make_blk(&b, &chk, false, false, kBlankUse))) return rc;
    g_h.blk = b;
    return LSM_OK;
}

int lsm_host_consumer_blanking(int32_t nblk, const int32_t *boxes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    LsmBlock b{};
    int rc = set_blanking(b, nblk, boxes);
    if (rc) return rc;
    g_c.nblk = b.nblk;
    for (int q = 0; q < LSM_MAX_BLANK; ++q) g_c.blank[q] = b.blank[q];
    return LSM_OK;
}

int lsm_host_invalidate(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_h.ready = false;   // the next lsm_host_setup uploads the grid again
    g_h.have_ucn = false;
    return LSM_OK;
}

int lsm_host_release(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    release();
    consumer_release();
    return LSM_OK;
}

int lsm_host_contra_velocity(const double *q, double *ucn_aos)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    // q(1:4, nodes): u, v, w are entries 2..4 of each group of four
    std::vector<double> tmp((size_t)n);
    double *uvw[3] = {g_h.f[1], g_h.f[2], g_h.f[3]};
    for (int c = 0; c < 3; ++c) {
        for (long l = 0; l < n; ++l) tmp[l] = q[4 * l + 1 + c];
        if ((rc = up(uvw[c], tmp.data(), n))) return rc;
    }
    if ((rc = lsm_contra_velocity(&g_h.blk, uvw[0], uvw[1], uvw[2], 1, g_h.ucn[0], g_h.ucn[1],
                                  g_h.ucn[2], nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    g_h.have_ucn = true;
    if (ucn_aos) {
        for (int c = 0; c < 3; ++c) {
            if ((rc = down(tmp.data(), g_h.ucn[c], n))) return rc;
            for (long l = 0; l < n; ++l) ucn_aos[3 * l + c] = tmp[l];
        }
    }
    return LSM_OK;
}

int lsm_host_weno3_rhs(const double *phi, double *rhs, const int32_t *mask, int mode)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!g_h.have_ucn) return fail(LSM_ERR_ARG, "contravariant velocity not computed (rhs_contra_j)");
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi, n)) || (rc = up(g_h.f[1], rhs, n))) return rc;
    if (mask && (rc = check_hip(hipMemcpy(g_h.mask, mask, n * sizeof(int32_t), hipMemcpyHostToDevice), "H2D")))
        return rc;
    if ((rc = lsm_weno3_rhs(&g_h.blk, g_h.f[0], g_h.ucn[0], g_h.ucn[1], g_h.ucn[2],
                            mask ? g_h.mask : nullptr, g_h.f[1], mode, nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    return down(rhs, g_h.f[1], n);
}

int lsm_host_upwind_rhs(const double *phi, double *rhs, const int32_t *mask)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!g_h.have_ucn) return fail(LSM_ERR_ARG, "contravariant velocity not computed (rhs_contra_j)");
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi, n)) || (rc = up(g_h.f[1], rhs, n))) return rc;
    if (mask && (rc = check_hip(hipMemcpy(g_h.mask, mask, n * sizeof(int32_t), hipMemcpyHostToDevice), "H2D")))
        return rc;
    if ((rc = lsm_upwind_rhs(&g_h.blk, g_h.f[0], g_h.ucn[0], g_h.ucn[1], g_h.ucn[2], mask ? g_h.mask : nullptr,
                             g_h.f[1], nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    return down(rhs, g_h.f[1], n);
}

int lsm_host_rk3(double *phi_np1, const double *phi_n, const int32_t *mask, double delti_lsm, int mode)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!g_h.have_ucn) return fail(LSM_ERR_ARG, "contravariant velocity not computed (rhs_contra_j)");
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi_n, n))) return rc;
    if (mask && (rc = check_hip(hipMemcpy(g_h.mask, mask, n * sizeof(int32_t), hipMemcpyHostToDevice), "H2D")))
        return rc;
    if ((rc = lsm_rk3_step(&g_h.blk, g_h.f[1], g_h.f[0], mask ? g_h.mask : nullptr, g_h.ucn[0],
                           g_h.ucn[1], g_h.ucn[2], delti_lsm, mode, nullptr, 0, nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    return down(phi_np1, g_h.f[1], n);
}

int lsm_host_levelset_step(double *phi, const double *phi_n, const double *q, const LsmReinitParams *p,
                           int numiter_lsm, double delti_lsm, int call_levelset, int call_reinit, int rn_freq,
                           int ntime, int mode, double vol[3])
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!phi || !phi_n || !p || (call_levelset && !q)) return fail(LSM_ERR_ARG, "NULL argument");
    const long n = g_h.n;
    if (!g_h.q && (rc = alloc((void **)&g_h.q, 4 * n * sizeof(double)))) return rc;
    const size_t need = lsm_workspace_bytes(&g_h.blk, LSM_OP_LEVELSET_STEP);
    if (need == 0) return fail(LSM_ERR_ARG, "lsm_workspace_bytes failed");
    if (g_h.lswork_bytes < need) {
        if (g_h.lswork) (void)hipFree(g_h.lswork);
        g_h.lswork = nullptr;
        g_h.lswork_bytes = 0;
        if ((rc = alloc(&g_h.lswork, need))) return rc;
        g_h.lswork_bytes = need;
    }
    // one upload of phi, phi_n and q; the whole call stays on the device
    if ((rc = up(g_h.f[0], phi, n)) || (rc = up(g_h.f[1], phi_n, n))) return rc;
    if (call_levelset && (rc = up(g_h.q, q, 4 * n))) return rc;
    if ((rc = lsm_levelset_step(&g_h.blk, p, g_h.f[0], g_h.f[1], g_h.q + 1, g_h.q + 2, g_h.q + 3, 4, numiter_lsm,
                                delti_lsm, call_levelset, call_reinit, rn_freq, ntime, mode, g_h.vol,
                                g_h.lswork, g_h.lswork_bytes, nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    if (vol && (rc = down(vol, g_h.vol, 3))) return rc;
    return down(phi, g_h.f[0], n);
}

int lsm_host_bcond(double *phi)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi, n))) return rc;
    if ((rc = lsm_bcond(&g_h.blk, g_h.f[0], nullptr))) return rc;
    if ((rc = sync())) return rc;
    return down(phi, g_h.f[0], n);
}

int lsm_host_normgrad(const double *d, const double *phi0, int order_bc, int eno_bc, double *ng)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], d, n)) || (rc = up(g_h.f[1], phi0, n)) || (rc = up(g_h.f[2], ng, n))) return rc;
    if ((rc = lsm_normgrad_eno2(&g_h.blk, g_h.f[0], g_h.f[1], order_bc, eno_bc, g_h.f[2], nullptr))) return rc;
    if ((rc = sync())) return rc;
    return down(ng, g_h.f[2], n);
}

int lsm_host_reinit_eno2(const double *phi0, double *phi_out, const LsmReinitParams *p, int32_t *iters)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi0, n))) return rc;
    if ((rc = lsm_reinit_sussman_eno2(&g_h.blk, p, g_h.f[0], g_h.f[1], g_h.iters, nullptr, 0, nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    if (iters && (rc = check_hip(hipMemcpy(iters, g_h.iters, sizeof(int32_t), hipMemcpyDeviceToHost), "D2H")))
        return rc;
    return down(phi_out, g_h.f[1], n);
}

int lsm_host_reinit_geometric(const double *phizero, double *phicorrected, int32_t *ids)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!g_h.blk.x) return fail(LSM_ERR_ARG, "coordinates were not passed to lsm_host_setup");
    if (!phizero || !phicorrected) return fail(LSM_ERR_ARG, "NULL argument");
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phizero, n)) || (rc = up(g_h.f[1], phicorrected, n))) return rc;
    if ((rc = lsm_reinit_geometric(&g_h.blk, g_h.f[0], g_h.f[1], ids ? g_h.mask : nullptr, nullptr, 0, nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    if ((rc = down(phicorrected, g_h.f[1], n))) return rc;
    if (ids) return check_hip(hipMemcpy(ids, g_h.mask, n * sizeof(int32_t), hipMemcpyDeviceToHost), "D2H");
    return LSM_OK;
}

int lsm_host_sign_ini(const double *phi0, double epslsm, int m_sphi0, double *sgn, double *ng)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi0, n)) || (rc = up(g_h.f[1], sgn, n)) || (rc = up(g_h.f[2], ng, n))) return rc;
    if ((rc = lsm_sign_ini(&g_h.blk, g_h.f[0], epslsm, m_sphi0, g_h.f[1], g_h.f[2], nullptr))) return rc;
    if ((rc = sync())) return rc;
    if ((rc = down(sgn, g_h.f[1], n))) return rc;
    return down(ng, g_h.f[2], n);
}

int lsm_host_rh_rn(const double *d, const double *sgn, double *rh)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], d, n)) || (rc = up(g_h.f[1], sgn, n)) || (rc = up(g_h.f[2], rh, n))) return rc;
    if ((rc = lsm_rh_rn(&g_h.blk, g_h.f[0], g_h.f[1], g_h.f[2], nullptr))) return rc;
    if ((rc = sync())) return rc;
    return down(rh, g_h.f[2], n);
}

int lsm_host_lambda_rn(const double *phi0, const double *ng0, const double *d, double dt, double eps,
                       double *lam)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi0, n)) || (rc = up(g_h.f[1], ng0, n)) || (rc = up(g_h.f[2], d, n)) ||
        (rc = up(g_h.f[3], lam, n)))
        return rc;
    if ((rc = lsm_lambda_rn(&g_h.blk, g_h.f[0], g_h.f[1], g_h.f[2], dt, eps, g_h.f[3], nullptr))) return rc;
    if ((rc = sync())) return rc;
    return down(lam, g_h.f[3], n);
}

int lsm_host_lambda_rn_simpson(const double *phi0, const double *ng0, const double *d, double dt, double eps,
                               double *lam)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi0, n)) || (rc = up(g_h.f[1], ng0, n)) || (rc = up(g_h.f[2], d, n)))
        return rc;
    if ((rc = lsm_lambda_rn_simpson(&g_h.blk, g_h.f[0], g_h.f[1], g_h.f[2], dt, eps, g_h.f[3], nullptr, 0,
                                    nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    return down(lam, g_h.f[3], n);
}

int lsm_host_reinit_godunov(const double *phi0, double *phi_out, const LsmReinitParams *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi0, n))) return rc;
    if ((rc = lsm_reinit_godunov(&g_h.blk, p, g_h.f[0], g_h.f[1], nullptr, 0, nullptr))) return rc;
    if ((rc = sync())) return rc;
    return down(phi_out, g_h.f[1], n);
}

int lsm_host_water_volume(const double *phi, double vol[3], int32_t *adv_mask)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = need_ready())) return rc;
    if (!g_h.blk.x) return fail(LSM_ERR_ARG, "coordinates were not passed to lsm_host_setup");
    const long n = g_h.n;
    if ((rc = up(g_h.f[0], phi, n))) return rc;
    if ((rc = lsm_water_volume(&g_h.blk, g_h.f[0], g_h.vol, adv_mask ? g_h.mask : nullptr, nullptr, 0,
                               nullptr)))
        return rc;
    if ((rc = sync())) return rc;
    if ((rc = check_hip(hipMemcpy(vol, g_h.vol, 3 * sizeof(double), hipMemcpyDeviceToHost), "D2H"))) return rc;
    if (adv_mask)
        return check_hip(hipMemcpy(adv_mask, g_h.mask, n * sizeof(int32_t), hipMemcpyDeviceToHost), "D2H");
    return LSM_OK;
}

int lsm_host_phi_gradient_eno2(const int32_t ext[9], int32_t phys_faces, double dc, double de, double dz,
                               const double *csi, const double *eta, const double *zet, const double *phi,
                               int order_bc, int eno_bc, double *grad)
{
    return consumer_gradient(ext, phys_faces, dc, de, dz, csi, eta, zet, phi, 3, order_bc, eno_bc, grad);
}

int lsm_host_h_gradient_eno2(const int32_t ext[9], int32_t phys_faces, double dc, double de, double dz,
                             const double *csi, const double *eta, const double *zet, const double *h,
                             int order_bc, int eno_bc, double *grad)
{
    return consumer_gradient(ext, phys_faces, dc, de, dz, csi, eta, zet, h, 2, order_bc, eno_bc, grad);
}

int lsm_host_calc_h(const int32_t ext[9], int32_t phys_faces, const double *z, const double *phi, double *h,
                    int32_t *missing)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!z || !phi || !h) return fail(LSM_ERR_ARG, "NULL argument");
    LsmBlock b;
    int rc = consumer_block(ext, phys_faces, 1.0, 1.0, 1.0, nullptr, nullptr, nullptr, &b);
    if (rc) return rc;
    const long n = g_c.n;
    if ((rc = up(g_c.f[0], z, n)) || (rc = up(g_c.f[1], phi, n))) return rc;
    if ((rc = lsm_calc_h(&b, g_c.f[0], g_c.f[1], g_c.f[2], g_c.cnt, nullptr))) return rc;
    if ((rc = sync())) return rc;
    if ((rc = down(h, g_c.f[2], n))) return rc;
    if (missing)
        return check_hip(hipMemcpy(missing, g_c.cnt, sizeof(int32_t), hipMemcpyDeviceToHost), "D2H");
    return LSM_OK;
}

}  // extern "C"
