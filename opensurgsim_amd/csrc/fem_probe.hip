// Device copy probe: a plain streaming copy of known size. It calibrates the FETCH_SIZE / WRITE_SIZE
// counters (bytes per counted KiB) and gives the achievable copy bandwidth of the device, beside the
// published HBM peak. It touches no simulation state.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/fem_gpu.h"

namespace {

__global__ __launch_bounds__(256) void probe_copy_kernel(int64_t n4, const float4* __restrict__ src,
                                                         float4* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void probe_fill_kernel(int64_t n4, float4* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n4) dst[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}

}  // namespace

extern "C" int fem_probe_copy(int device, int64_t bytes, int reps, double* ms_per_copy) {
    if (bytes < 16 || reps < 1 || !ms_per_copy) return FEM_ERR_INVALID;
    const int64_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FEM_ERR_HIP;
    const dim3 grid((unsigned)((n4 + 255) / 256)), block(256);
    if (hipSetDevice(device) != hipSuccess) return FEM_ERR_HIP;
    if (hipMalloc(&a, n4 * 16) != hipSuccess || hipMalloc(&b, n4 * 16) != hipSuccess) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    hipLaunchKernelGGL(probe_fill_kernel, grid, block, 0, 0, n4, a);
    hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, 0, n4, (const float4*)a, b);  // warm-up
    if (hipEventRecord(e0, 0) != hipSuccess) goto done;
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, 0, n4, (const float4*)a, b);
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
    {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto done;
        *ms_per_copy = t / reps;
        rc = FEM_OK;
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
