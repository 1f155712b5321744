// Deterministic wavefront / workgroup / grid reductions shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fem_kernels.hpp"

namespace femgpu {

// ------------------------------------------------------------------------------------------
// Reductions (deterministic: fixed lane/wave order).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums of two values over the block with a single barrier; results valid in thread 0.
// lds must hold 2 * blockDim/64 doubles.
__device__ __forceinline__ void block_sum2(double& v0, double& v1, double* lds) {
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[w] = v0;
        lds[nw + w] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        v0 = wave_sum(threadIdx.x < (unsigned)nw ? lds[threadIdx.x] : 0.0);
        v1 = wave_sum(threadIdx.x < (unsigned)nw ? lds[nw + threadIdx.x] : 0.0);
    }
}

// Sum over the block; result valid in thread 0. lds must hold blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    int nw = blockDim.x >> 6;
    double t = 0.0;
    if (threadIdx.x < 64) {
        t = (threadIdx.x < (unsigned)nw) ? lds[threadIdx.x] : 0.0;
        t = wave_sum(t);
    }
    __syncthreads();
    return t;
}

// Two sums of kCgBlocks partials each, formed by wave 0 in a fixed order and broadcast through LDS.
__device__ __forceinline__ void reduce_partials2(const double* pa, const double* pb, double& a, double& b,
                                                 double* lds) {
    if (threadIdx.x < 64) {
        double va = 0.0, vb = 0.0;
        for (int i = threadIdx.x; i < kCgBlocks; i += 64) {
            va += pa[i];
            if (pb) vb += pb[i];
        }
        va = wave_sum(va);
        vb = wave_sum(vb);
        if (threadIdx.x == 0) {
            lds[62] = va;
            lds[63] = vb;
        }
    }
    __syncthreads();
    a = lds[62];
    b = lds[63];
    __syncthreads();
}

__device__ __forceinline__ double reduce_partials(const double* part, double* lds) {
    double a, b;
    reduce_partials2(part, nullptr, a, b, lds);
    return a;
}

}  // namespace femgpu
