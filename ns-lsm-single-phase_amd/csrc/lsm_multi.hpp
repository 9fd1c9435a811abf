// Multi-instance execution of the single-block entry points.
//
// A kernel is described by a struct K with
//     using Args = ...;                       // everything it reads, by value
//     static constexpr int kThreads = ...;    // threads per workgroup
//     __device__ static void run(const Args &a, int bid, int nb);
//     static void one(unsigned grid, dim3 block, hipStream_t s, const Args &a);
// where bid / nb replace blockIdx.x / gridDim.x and `one` launches the
// kernel's own __global__ wrapper (which calls run with blockIdx.x and
// gridDim.x) on one block.  klaunch<K> calls `one`, or, inside a recording
// scope (lsm_multi_begin .. lsm_multi_end), appends the launch to the current
// instance's list.  lsm_multi_end merges position p of every
// instance into one launch of kmulti<K> with a 2-D grid: blockIdx.y is the
// instance, blockIdx.x the workgroup within it (workgroups past the
// instance's own count return at once), the instance's Args read from an
// array in device memory.  The x extent is a multiple of 8, so workgroup
// bid of every instance runs on XCD bid % 8 as in its own launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "lsm_internal.hpp"

namespace lsm {

template <class K>
__global__ __launch_bounds__(K::kThreads) void kmulti(const typename K::Args *__restrict__ args,
                                                      const int *__restrict__ grids)
{
    const int b = (int)blockIdx.y;
    const int nb = grids[b];
    if ((int)blockIdx.x >= nb) return;
    const typename K::Args a = args[b];
    K::run(a, (int)blockIdx.x, nb);
}

// launches kmulti<K> over n instances whose Args (n of them) and grid sizes
// are in device memory
using MultiLaunchFn = int (*)(unsigned gx, unsigned n, dim3 block, hipStream_t s, const void *args,
                              const int *grids);

template <class K>
int multi_launch(unsigned gx, unsigned n, dim3 block, hipStream_t s, const void *args, const int *grids)
{
    kmulti<K><<<dim3(gx, n, 1), block, 0, s>>>((const typename K::Args *)args, grids);
    return check_hip(hipGetLastError(), "multi-instance launch");
}

// whether the calling thread is inside lsm_multi_begin .. lsm_multi_end
bool recording();
// appends one launch to the current instance (copies the arguments)
int record_launch(MultiLaunchFn fn, unsigned grid, dim3 block, const void *args, size_t args_size,
                  const char *what);

template <class K>
int klaunch(unsigned grid, dim3 block, hipStream_t s, const typename K::Args &a, const char *what)
{
    if (recording()) return record_launch(&multi_launch<K>, grid, block, &a, sizeof(a), what);
    if (grid == 0) return LSM_OK;
    K::one(grid, block, s, a);
    return check_hip(hipGetLastError(), what);
}

// Entry points that can be recorded declare it at their top; make_blk
// rejects every other single-block entry point while recording.
struct Recordable {
    Recordable();
    ~Recordable();
};
bool recordable_here();

}  // namespace lsm
