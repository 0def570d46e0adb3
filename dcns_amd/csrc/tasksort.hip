// tasksort.hip -- orders the recursion's scan tasks by row count (longest
// first) so that the 64 lanes of a k_scan_general wave, one task each, run
// loops of similar length.  The tasks stay where they are; the scan kernel
// reads them through the sorted index.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace dcns {
namespace {

__global__ void k_task_keys(const ScanTask *__restrict__ tasks, const int32_t *__restrict__ n_dev,
                            uint32_t *__restrict__ keys, int32_t *__restrict__ idx, int32_t n_max) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_max) return;
    const int32_t n = *n_dev;
    // tasks past the live count sort last (key 0) and are never read as live ones
    keys[i] = i < n ? (uint32_t)max(tasks[i].b - tasks[i].a, 0) : 0u;
    idx[i] = i;
}

}  // namespace

size_t sort_tasks_scratch_bytes(int32_t n_max) {
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                 (const int32_t *)nullptr, (int32_t *)nullptr, n_max, 0, 25);
    return bytes + 4 * (size_t)n_max * 4 + 1024;
}

hipError_t launch_sort_tasks(hipStream_t st, const ScanTask *tasks, const int32_t *n_dev, int32_t n_max,
                             void *scratch, size_t scratch_bytes, int32_t *order) {
    if (n_max <= 0) return hipSuccess;
    // scratch: keys in / out, values in, then the sort's temporary storage
    char *p = (char *)scratch;
    uint32_t *kin = (uint32_t *)p;
    uint32_t *kout = kin + n_max;
    int32_t *vin = (int32_t *)(kout + n_max);
    char *tmp = (char *)(vin + n_max);
    size_t tmp_bytes = scratch_bytes - 3 * (size_t)n_max * 4;
    hipLaunchKernelGGL(k_task_keys, dim3((n_max + 255) / 256), dim3(256), 0, st, tasks, n_dev, kin, vin, n_max);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // row counts stay below 2^25 (sequences are at most 16 M bases)
    return hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, kin, kout, vin, order, n_max, 0, 25, st);
}

}  // namespace dcns
