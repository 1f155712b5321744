// Multi-column FP64 conjugate gradients (see fem_mcg.hpp).
//
// Every column follows Eigen's ConjugateGradient with the Jacobi preconditioner (the recurrence of
// Math/LinearSparseSolveAndInverse.cpp:99-129 over Eigen 3.3's conjugate_gradient): x = 0, r = b, stop when
// |r|^2 < max(tol^2 |b|^2, realmin) or at max_iter iterations. The columns of a batch share the operator
// reads: one wavefront per block row, lanes = (64 / KW) block slots x KW columns, right-hand sides and work
// vectors interleaved dof-major ([dof][KW]) so that the KW values of a gathered column are contiguous.
// One iteration is three launches: q = S p (+ p.q), the x, r, z updates (+ r.r, r.z) and the convergence
// test with the new direction. Sums are per column over fixed lane / wave / block orders.
#include "fem_mcg.hpp"
#include "fem_kernels.hpp"
#include "fem_reduce.hpp"

#include <cfloat>
#include <mutex>
#include <stdexcept>
#include <string>

namespace femgpu {

// The allocation-free entry point's workspace (uninitialised; every kernel that reads a part of it writes
// that part first in the same call). Host code never passes `scratch` itself: its name on the host is a
// stand-in, not the device address, which is taken per device below.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void* noalloc_scratch_address(int device) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[device] = ptr;
    }
    return addr[device];
}

McgLayout mcg_layout(size_t n) {
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    McgLayout L;
    size_t o = 0;
    const size_t vec = up(n * 8 * sizeof(double));
    L.b = o, o += vec;
    L.x = o, o += vec;
    L.r = o, o += vec;
    L.z = o, o += vec;
    L.p = o, o += vec;
    L.q = o, o += vec;
    L.dinv = o, o += up(n * sizeof(double));
    L.partial = o, o += up((size_t)3 * kCgBlocks * 8 * sizeof(double));
    L.col = o, o += up(8 * sizeof(McgColumn));
    L.total = o;
    return L;
}

namespace {

// Block sums per column of values held by the threads of column j = lane % KW; valid in threads < KW.
// lds: (blockDim / 64) * KW doubles.
template <int KW>
__device__ __forceinline__ double col_block_sum(double v, double* lds) {
#pragma unroll
    for (int o = KW; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane < KW) lds[w * KW + lane] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < KW)
        for (int q = 0; q < nw; ++q) t += lds[q * KW + threadIdx.x];
    __syncthreads();
    return t;
}

// Per-column totals of kCgBlocks partials laid out [kCgBlocks][KW], formed by every thread of the block in a
// fixed order (each adds a strided share, then the per-column block sum); out[0..KW) (LDS) is valid for every
// thread after the call. lds: (blockDim / 64) * KW doubles.
template <int KW>
__device__ __forceinline__ void col_reduce(const double* part, double* out, double* lds) {
    const int j = threadIdx.x % KW;
    double v = 0.0;
    for (int b = threadIdx.x / KW; b < kCgBlocks; b += blockDim.x / KW) v += part[b * KW + j];
    v = col_block_sum<KW>(v, lds);
    if (threadIdx.x < KW) out[threadIdx.x] = v;
    __syncthreads();
}

// x = 0, r = b, z = p = D^-1 r with D^-1 the Jacobi inverse (1 / a_ii, 1 where a_ii = 0); partial |b|^2, r.z.
template <int KW>
__global__ __launch_bounds__(kCgThreads) void mcg_init_kernel(McgArgs a) {
    __shared__ double lds[(kCgThreads / 64) * KW];
    const int64_t total = 3 * (int64_t)a.num_nodes * KW;
    double bb = 0.0, rz = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / KW;
        const int I = (int)(i / 3), d = (int)(i - 3 * (int64_t)I);
        const int db = a.diagblk[I];
        const double diag = db >= 0 ? a.S[9 * (int64_t)db + 4 * d] : 0.0;
        const double di = diag != 0.0 ? 1.0 / diag : 1.0;
        if (t % KW == 0) a.dinv[i] = di;
        const double bv = a.b[t];
        const double zv = di * bv;
        a.x[t] = 0.0;
        a.r[t] = bv;
        a.z[t] = zv;
        a.p[t] = zv;
        bb += bv * bv;
        rz += bv * zv;
    }
    const double sb = col_block_sum<KW>(bb, lds);
    const double sz = col_block_sum<KW>(rz, lds);
    if (threadIdx.x < KW) {
        a.partial[blockIdx.x * KW + threadIdx.x] = sb;
        a.partial[(kCgBlocks + blockIdx.x) * KW + threadIdx.x] = sz;
    }
}

// Column set-up: |b|^2, the threshold, r.z, and the columns that finish before iterating.
template <int KW>
__global__ __launch_bounds__(256) void mcg_setup_kernel(McgArgs a) {
    __shared__ double s[16];
    __shared__ double lds[4 * KW];
    col_reduce<KW>(a.partial, s, lds);
    col_reduce<KW>(a.partial + kCgBlocks * KW, s + 8, lds);
    if (threadIdx.x < KW) {
        const int j = threadIdx.x;
        McgColumn c;
        c.bb = s[j];
        c.rz[0] = s[8 + j];
        c.rz[1] = 0.0;
        c.threshold = fmax(a.tol * a.tol * c.bb, DBL_MIN);
        c.rr = c.bb;
        c.iters = 0;
        c.pad = 0;
        c.converged = 1;
        c.fin = kMcgRunning;
        if (c.bb == 0.0 || c.bb < c.threshold) c.fin = -1;  // zero rhs, or x = 0 already meets the test
        else if (a.max_iter <= 0) {
            c.fin = -1;
            c.converged = 0;
        }
        a.col[j] = c;
    }
}

// q = S p for the running columns, partial p.q.
template <int KW>
__global__ __launch_bounds__(kCgThreads) void mcg_spmv_kernel(McgArgs a, int k) {
    __shared__ double lds[(kCgThreads / 64) * KW];
    constexpr int NS = 64 / KW;
    const int lane = threadIdx.x & 63, j = lane % KW, slot = lane / KW;
    const bool act = a.col[j].fin >= k;
    const int wave = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)(((int64_t)gridDim.x * blockDim.x) >> 6);
    double pq = 0.0;
    for (int row = wave; row < a.num_nodes; row += nwaves) {
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        if (act) {
            const int b1 = a.rowptr[row + 1];
            int blk = a.rowptr[row] + slot;
            int jn = blk < b1 ? a.colidx[blk] : 0;
            for (; blk < b1; blk += NS) {
                const double* Sb = a.S + 9 * (int64_t)blk;
                const int64_t cb = 3 * (int64_t)jn * KW + j;
                if (blk + NS < b1) jn = a.colidx[blk + NS];  // the next block's column goes out early
                const double p0 = a.p[cb], p1 = a.p[cb + KW], p2 = a.p[cb + 2 * KW];
                acc0 += Sb[0] * p0 + Sb[1] * p1 + Sb[2] * p2;
                acc1 += Sb[3] * p0 + Sb[4] * p1 + Sb[5] * p2;
                acc2 += Sb[6] * p0 + Sb[7] * p1 + Sb[8] * p2;
            }
        }
#pragma unroll
        for (int o = KW; o < 64; o <<= 1) {
            acc0 += __shfl_xor(acc0, o, 64);
            acc1 += __shfl_xor(acc1, o, 64);
            acc2 += __shfl_xor(acc2, o, 64);
        }
        if (act && slot == 0) {
            const int64_t r0 = 3 * (int64_t)row * KW + j;
            a.q[r0] = acc0;
            a.q[r0 + KW] = acc1;
            a.q[r0 + 2 * KW] = acc2;
            pq += a.p[r0] * acc0 + a.p[r0 + KW] * acc1 + a.p[r0 + 2 * KW] * acc2;
        }
    }
    const double s = col_block_sum<KW>(pq, lds);
    if (threadIdx.x < KW) a.partial[blockIdx.x * KW + threadIdx.x] = s;
}

// alpha = r.z / p.q; x += alpha p; r -= alpha q; z = D^-1 r; partial r.r, r.z.
template <int KW>
__global__ __launch_bounds__(kCgThreads) void mcg_vector_kernel(McgArgs a, int k) {
    __shared__ double lds[(kCgThreads / 64) * KW];
    __shared__ double alpha[8];
    col_reduce<KW>(a.partial, alpha, lds);
    if (threadIdx.x < KW) alpha[threadIdx.x] = a.col[threadIdx.x].rz[k & 1] / alpha[threadIdx.x];
    __syncthreads();
    const int j = threadIdx.x % KW;
    const bool act = a.col[j].fin >= k;
    const double al = alpha[j];
    double rr = 0.0, rz = 0.0;
    if (act) {
        const int64_t total = 3 * (int64_t)a.num_nodes * KW;
        for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
             t += (int64_t)gridDim.x * blockDim.x) {
            a.x[t] += al * a.p[t];
            const double rv = a.r[t] - al * a.q[t];
            const double zv = a.dinv[t / KW] * rv;
            a.r[t] = rv;
            a.z[t] = zv;
            rr += rv * rv;
            rz += rv * zv;
        }
    }
    const double s0 = col_block_sum<KW>(rr, lds);
    const double s1 = col_block_sum<KW>(rz, lds);
    if (threadIdx.x < KW) {
        a.partial[(kCgBlocks + blockIdx.x) * KW + threadIdx.x] = s0;
        a.partial[(2 * kCgBlocks + blockIdx.x) * KW + threadIdx.x] = s1;
    }
}

// Convergence test of iteration k and p = z + beta p for the columns that go on. Block 0 records the
// column state; the other blocks decide from the same reduced sums (fin is only read as >= k, which holds
// before and after block 0 writes k; r.z is double-buffered by parity).
template <int KW>
__global__ __launch_bounds__(kCgThreads) void mcg_direction_kernel(McgArgs a, int k) {
    __shared__ double rr[8], rzn[8], beta[8];
    __shared__ int go[8];
    __shared__ double lds[(kCgThreads / 64) * KW];
    col_reduce<KW>(a.partial + kCgBlocks * KW, rr, lds);
    col_reduce<KW>(a.partial + 2 * kCgBlocks * KW, rzn, lds);
    if (threadIdx.x < KW) {
        const int j = threadIdx.x;
        const int fin = a.col[j].fin;
        const double rz_old = a.col[j].rz[k & 1], thr = a.col[j].threshold;
        int g = 0;
        double be = 0.0;
        if (fin >= k) {
            if (rr[j] < thr) {
                if (blockIdx.x == 0) {
                    a.col[j].iters = k;
                    a.col[j].converged = 1;
                }
            } else if (k + 1 >= a.max_iter) {
                if (blockIdx.x == 0) {
                    a.col[j].iters = k + 1;
                    a.col[j].converged = 0;
                }
            } else {
                g = 1;
                be = rzn[j] / rz_old;
                if (blockIdx.x == 0) a.col[j].rz[(k + 1) & 1] = rzn[j];
            }
            if (blockIdx.x == 0) {
                a.col[j].rr = rr[j];
                if (!g) a.col[j].fin = k;
            }
        }
        go[j] = g;
        beta[j] = be;
    }
    __syncthreads();
    const int j = threadIdx.x % KW;
    if (!go[j]) return;
    const double be = beta[j];
    const int64_t total = 3 * (int64_t)a.num_nodes * KW;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
        a.p[t] = a.z[t] + be * a.p[t];
}

template <int KW>
void init_kw(const McgArgs& a, hipStream_t s) {
    launch_k(mcg_init_kernel<KW>, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a);
    launch_k(mcg_setup_kernel<KW>, dim3(1), dim3(256), 0, s, a);
}

template <int KW>
void iteration_kw(const McgArgs& a, int k, hipStream_t s) {
    launch_k(mcg_spmv_kernel<KW>, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a, k);
    launch_k(mcg_vector_kernel<KW>, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a, k);
    launch_k(mcg_direction_kernel<KW>, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a, k);
}

// dst[i kw + j] = b[j n + i] (0 on boundary rows and for padding columns j >= cnt)
__global__ __launch_bounds__(256) void mcg_gather_kernel(int64_t n, int kw, int cnt, const double* __restrict__ b,
                                                         const uint8_t* __restrict__ mask, double* __restrict__ dst) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * kw) return;
    const int64_t i = t / kw;
    const int j = (int)(t - i * kw);
    dst[t] = (j < cnt && !mask[i]) ? b[(int64_t)j * n + i] : 0.0;
}

// out[j n + i] = x[i kw + j] (0 on boundary rows), j < cnt
__global__ __launch_bounds__(256) void mcg_scatter_kernel(int64_t n, int kw, int cnt, const double* __restrict__ x,
                                                          const uint8_t* __restrict__ mask, double* __restrict__ out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * cnt) return;
    const int j = (int)(t / n);
    const int64_t i = t - (int64_t)j * n;
    out[t] = mask[i] ? 0.0 : x[i * kw + j];
}

// dst = src with boundary rows 0
__global__ __launch_bounds__(256) void masked_copy_kernel(int64_t n, const double* __restrict__ src,
                                                          const uint8_t* __restrict__ mask, double* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = mask[i] ? 0.0 : src[i];
}

// dst[idx[k]] = val[k] (distinct indices), and dst[k] = src[idx[k]]: the few non-zeros of a localized
// constraint's H^T columns and the solution entries at them
__global__ __launch_bounds__(64) void scatter_values_kernel(int cnt, const int64_t* __restrict__ idx,
                                                            const double* __restrict__ val, double* __restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < cnt) dst[idx[k]] = val[k];
}
__global__ __launch_bounds__(64) void gather_values_kernel(int cnt, const int64_t* __restrict__ idx,
                                                           const double* __restrict__ src, double* __restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < cnt) dst[k] = src[idx[k]];
}

}  // namespace

void launch_scatter_values(int cnt, const int64_t* idx, const double* val, double* dst, hipStream_t s) {
    if (cnt > 0) launch_k(scatter_values_kernel, dim3((unsigned)((cnt + 63) / 64)), dim3(64), 0, s, cnt, idx, val, dst);
}
void launch_gather_values(int cnt, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
    if (cnt > 0) launch_k(gather_values_kernel, dim3((unsigned)((cnt + 63) / 64)), dim3(64), 0, s, cnt, idx, src, dst);
}

void launch_mcg_gather(int64_t n, int kw, int cnt, const double* b, const uint8_t* mask, double* dst, hipStream_t s) {
    launch_k(mcg_gather_kernel, dim3((unsigned)((n * kw + 255) / 256)), dim3(256), 0, s, n, kw, cnt, b, mask, dst);
}
void launch_mcg_scatter(int64_t n, int kw, int cnt, const double* x, const uint8_t* mask, double* out, hipStream_t s) {
    launch_k(mcg_scatter_kernel, dim3((unsigned)((n * cnt + 255) / 256)), dim3(256), 0, s, n, kw, cnt, x, mask, out);
}
void launch_masked_copy(int64_t n, const double* src, const uint8_t* mask, double* dst, hipStream_t s) {
    launch_k(masked_copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, mask, dst);
}

void launch_mcg_init(const McgArgs& a, hipStream_t s) {
    switch (a.kw) {
        case 1: init_kw<1>(a, s); break;
        case 2: init_kw<2>(a, s); break;
        case 4: init_kw<4>(a, s); break;
        default: init_kw<8>(a, s); break;
    }
}

void launch_mcg_iteration(const McgArgs& a, int k, hipStream_t s) {
    switch (a.kw) {
        case 1: iteration_kw<1>(a, k, s); break;
        case 2: iteration_kw<2>(a, k, s); break;
        case 4: iteration_kw<4>(a, k, s); break;
        default: iteration_kw<8>(a, k, s); break;
    }
}

}  // namespace femgpu
