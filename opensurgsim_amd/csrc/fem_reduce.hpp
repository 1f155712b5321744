// Deterministic wavefront / workgroup / grid reductions shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fem_kernels.hpp"

namespace femgpu {

// ------------------------------------------------------------------------------------------
// Reductions (deterministic: fixed lane/wave order).
// ------------------------------------------------------------------------------------------
// Butterfly over the wavefront by shuffles (ds_bpermute); correct with any set of active lanes that is
// closed under the butterfly, result in every lane.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// DPP lane exchange (v_mov_b32 with a data-parallel-primitive modifier): every lane reads a fixed other
// lane of its wavefront without going through LDS. Lanes of rows masked off by ROW get 0.
//   0xB1 quad_perm[1,0,3,2] (lane ^ 1)       0x4E quad_perm[2,3,0,1] (lane ^ 2)
//   0x141 row_half_mirror (i <- 7 - i in 8)  0x140 row_mirror (i <- 15 - i in 16)
//   0x142 row_bcast:15 (rows 1, 3 <- last lane of the row before)   0x143 row_bcast:31 (rows 2, 3 <- lane 31)
template <int CTRL, int ROW = 0xF>
__device__ __forceinline__ float dpp_mov(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW, 0xF, false));
}
template <int CTRL, int ROW = 0xF>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Sum over aligned groups of G lanes, in every lane of the group (G = 4, 8, 16 by DPP, others by
// shuffles). Every lane of the group must be active.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
    if constexpr (G == 4 || G == 8 || G == 16) {
        v += dpp_mov<0xB1>(v);
        v += dpp_mov<0x4E>(v);
        if constexpr (G >= 8) v += dpp_mov<0x141>(v);
        if constexpr (G == 16) v += dpp_mov<0x140>(v);
    } else {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    }
    return v;
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a wavefront by DPP (rows of 16, then row broadcasts; lane 63 ends with the
// total), returned uniformly. All 64 lanes must be active.
__device__ __forceinline__ double wave_total(double v) {
    v = group_sum<16>(v);
    v += dpp_mov<0x142, 0xA>(v);
    v += dpp_mov<0x143, 0xC>(v);
    return read_lane(v, 63);
}

// Sums of two values over the block with a single barrier; results valid in thread 0. Every thread of the
// block calls it. lds must hold 2 * blockDim/64 doubles.
__device__ __forceinline__ void block_sum2(double& v0, double& v1, double* lds) {
    v0 = wave_total(v0);
    v1 = wave_total(v1);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[w] = v0;
        lds[nw + w] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        v0 = wave_total(threadIdx.x < (unsigned)nw ? lds[threadIdx.x] : 0.0);
        v1 = wave_total(threadIdx.x < (unsigned)nw ? lds[nw + threadIdx.x] : 0.0);
    }
}

// Sum over the block; result valid in thread 0. Every thread of the block calls it. lds must hold
// blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_total(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    int nw = blockDim.x >> 6;
    double t = 0.0;
    if (threadIdx.x < 64) {
        t = (threadIdx.x < (unsigned)nw) ? lds[threadIdx.x] : 0.0;
        t = wave_total(t);
    }
    __syncthreads();
    return t;
}

// Three sums of kCgBlocks partials each (pb, pc may be null), formed by wave 0 in a fixed order (lane l adds
// entries l, l + 64, ... in that order, then the wave total) and broadcast through LDS (lds[61..63]). Every load
// of a lane goes out before the first add: the partials come from workgroups on every XCD, so each load misses
// this XCD's L2, and a loop that waits for one load at a time pays that latency kCgBlocks / 64 times.
__device__ __forceinline__ void reduce_partials3(const double* pa, const double* pb, const double* pc, double& a,
                                                 double& b, double& c, double* lds) {
    constexpr int kPer = kCgBlocks / 64;
    if (threadIdx.x < 64) {
        double ta[kPer], tb[kPer], tc[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = threadIdx.x + 64 * q;
            ta[q] = pa[i];
            tb[q] = pb ? pb[i] : 0.0;
            tc[q] = pc ? pc[i] : 0.0;
        }
        double va = 0.0, vb = 0.0, vc = 0.0;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            va += ta[q];
            vb += tb[q];
            vc += tc[q];
        }
        va = wave_total(va);
        vb = wave_total(vb);
        vc = wave_total(vc);
        if (threadIdx.x == 0) {
            lds[61] = va;
            lds[62] = vb;
            lds[63] = vc;
        }
    }
    __syncthreads();
    a = lds[61];
    b = lds[62];
    c = lds[63];
    __syncthreads();
}

// Two sums of kCgBlocks partials each (pb may be null), as reduce_partials3.
__device__ __forceinline__ void reduce_partials2(const double* pa, const double* pb, double& a, double& b,
                                                 double* lds) {
    double c;
    reduce_partials3(pa, pb, nullptr, a, b, c, lds);
}

__device__ __forceinline__ double reduce_partials(const double* part, double* lds) {
    double a, b;
    reduce_partials2(part, nullptr, a, b, lds);
    return a;
}

}  // namespace femgpu
