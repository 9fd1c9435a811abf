// Recording and merging of single-block launch sequences (lsm_multi_*).
//
// Between lsm_multi_begin and lsm_multi_end the recordable entry points do
// not launch: each kernel launch is appended, with a copy of its arguments,
// to the current instance (lsm_multi_instance starts the next one).
// lsm_multi_end requires every instance to have recorded the same sequence
// of kernels and block shapes, uploads all arguments and per-instance grid
// sizes with one host-to-device copy, and issues one 2-D launch per
// position of the sequence (kmulti<K>, lsm_multi.hpp).  Instances are
// therefore advanced together, stage by stage, in one launch per kernel.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "lsm_internal.hpp"
#include "lsm_multi.hpp"

namespace lsm {

namespace {

struct Rec {
    MultiLaunchFn fn;
    unsigned grid;
    dim3 block;
    size_t off, size;   // argument bytes in Recorder::bytes
    const char *what;
};

struct Recorder {
    std::vector<std::vector<Rec>> inst;
    std::vector<unsigned char> bytes;
};

thread_local Recorder *g_rec = nullptr;
thread_local int g_recordable = 0;

// pinned staging of the arguments, one per (device, stream); `ev` marks the
// end of the last copy out of it, which must complete before it is refilled
struct Staging {
    void *host = nullptr;
    size_t size = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

std::mutex g_stage_mu;
std::map<std::pair<int, uintptr_t>, Staging> g_stage;

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

bool recording() { return g_rec != nullptr; }

Recordable::Recordable() { ++g_recordable; }
Recordable::~Recordable() { --g_recordable; }
bool recordable_here() { return g_recordable > 0; }

int record_launch(MultiLaunchFn fn, unsigned grid, dim3 block, const void *args, size_t args_size,
                  const char *what)
{
    if (g_rec->inst.empty()) return fail(LSM_ERR_ARG, "lsm_multi_instance has not been called");
    Rec r{fn, grid, block, g_rec->bytes.size(), args_size, what};
    const unsigned char *p = (const unsigned char *)args;
    g_rec->bytes.insert(g_rec->bytes.end(), p, p + args_size);
    g_rec->inst.back().push_back(r);
    return LSM_OK;
}

}  // namespace lsm

using namespace lsm;

extern "C" {

int lsm_multi_begin(void)
{
    if (g_rec) return fail(LSM_ERR_ARG, "lsm_multi_begin: already recording on this thread");
    g_rec = new Recorder;
    return LSM_OK;
}

int lsm_multi_instance(void)
{
    if (!g_rec) return fail(LSM_ERR_ARG, "lsm_multi_instance outside lsm_multi_begin .. lsm_multi_end");
    g_rec->inst.emplace_back();
    return LSM_OK;
}

int lsm_multi_cancel(void)
{
    delete g_rec;
    g_rec = nullptr;
    return LSM_OK;
}

int lsm_multi_end(void *stream)
{
    std::unique_ptr<Recorder> R(g_rec);
    g_rec = nullptr;
    if (!R) return fail(LSM_ERR_ARG, "lsm_multi_end without lsm_multi_begin");
    const size_t n = R->inst.size();
    if (n == 0) return LSM_OK;
    if (n > 65535) return fail(LSM_ERR_ARG, "%zu instances: at most 65535", n);
    const size_t np = R->inst[0].size();
    for (size_t b = 1; b < n; ++b)
        if (R->inst[b].size() != np)
            return fail(LSM_ERR_ARG, "instance %zu recorded %zu launches, instance 0 recorded %zu", b,
                        R->inst[b].size(), np);
    // per position: the grid sizes (n ints), then the n argument blocks
    std::vector<size_t> goff(np), aoff(np);
    std::vector<unsigned> gx(np);
    size_t total = 0;
    for (size_t p = 0; p < np; ++p) {
        const Rec &r0 = R->inst[0][p];
        unsigned mx = 0;
        for (size_t b = 0; b < n; ++b) {
            const Rec &r = R->inst[b][p];
            if (r.fn != r0.fn || r.size != r0.size || r.block.x != r0.block.x || r.block.y != r0.block.y ||
                r.block.z != r0.block.z)
                return fail(LSM_ERR_ARG, "launch %zu differs between instances 0 (%s) and %zu (%s)", p, r0.what, b,
                            r.what);
            mx = std::max(mx, r.grid);
        }
        gx[p] = (mx + 7) / 8 * 8;
        goff[p] = total;
        total = al256(total + n * sizeof(int));
        aoff[p] = total;
        total = al256(total + n * r0.size);
    }
    if (np == 0) return LSM_OK;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, rc;
    if ((rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))) return rc;
    void *dbuf = nullptr;
    if ((rc = cached_buffer(kBufMulti, total, s, &dbuf))) return rc;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging &S = g_stage[std::make_pair(dev, (uintptr_t)s)];
    if (S.pending && (rc = check_hip(hipEventSynchronize(S.ev), "hipEventSynchronize"))) return rc;
    S.pending = false;
    if (!S.ev && (rc = check_hip(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming), "hipEventCreate")))
        return rc;
    if (S.size < total) {
        if (S.host) (void)hipHostFree(S.host);
        S.host = nullptr;
        S.size = 0;
        if ((rc = check_hip(hipHostMalloc(&S.host, total + total / 4, hipHostMallocDefault), "hipHostMalloc")))
            return rc;
        S.size = total + total / 4;
    }
    unsigned char *h = (unsigned char *)S.host;
    for (size_t p = 0; p < np; ++p) {
        int *g = (int *)(h + goff[p]);
        const size_t sz = R->inst[0][p].size;
        for (size_t b = 0; b < n; ++b) {
            const Rec &r = R->inst[b][p];
            g[b] = (int)r.grid;
            memcpy(h + aoff[p] + b * sz, R->bytes.data() + r.off, sz);
        }
    }
    if ((rc = check_hip(hipMemcpyAsync(dbuf, S.host, total, hipMemcpyHostToDevice, s), "hipMemcpyAsync")))
        return rc;
    if ((rc = check_hip(hipEventRecord(S.ev, s), "hipEventRecord"))) return rc;
    S.pending = true;
    const unsigned char *d = (const unsigned char *)dbuf;
    for (size_t p = 0; p < np; ++p) {
        if (gx[p] == 0) continue;
        const Rec &r0 = R->inst[0][p];
        if ((rc = r0.fn(gx[p], (unsigned)n, r0.block, s, d + aoff[p], (const int *)(d + goff[p])))) return rc;
    }
    return LSM_OK;
}

}  // extern "C"
