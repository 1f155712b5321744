// Banded LU without pivoting in FP64 (see fem_lu.hpp).
//
// Factorization: right-looking, blocked by panels of kLuPanel columns. For panel [k0, k1):
//   1. one workgroup factors the panel column by column: multipliers l(i, c) = a(i, c) / a(c, c) for the rows
//      below the diagonal within the band, and the rank-1 updates restricted to the panel's columns (the L21
//      block, rows up to k1 - 1 + bw) and the panel's rows (the U12 block, columns up to k1 - 1 + bw);
//   2. the trailing block a(i, j), i, j in [k1, k1 + bw), takes -= L21 U12, by 64 x 64 output tiles staged
//      through LDS (the entries of L21 and U12 outside the band are zero).
// The same elimination as the column-by-column form of the restated reference solver, with the panel's updates
// of the trailing block summed before they are applied, so results agree to rounding.
//
// Substitution: one workgroup per right-hand side. Rows go in blocks of 64: the block's products with the
// already solved entries (up to bw of them per row) are summed by 16 lanes per row, then one wavefront solves
// the 64 x 64 triangle, staged in LDS, with the solved entries broadcast lane to lane.
#include "fem_lu.hpp"
#include "fem_kernels.hpp"

#include <algorithm>
#include <stdexcept>

namespace femgpu {

namespace {

constexpr int kLuPanel = 32;
constexpr int kTile = 64;

__device__ __forceinline__ double& at(double* ab, int64_t ld, int bw, int64_t i, int64_t j) {
    return ab[i * ld + (j - i + bw)];
}
__device__ __forceinline__ double atc(const double* ab, int64_t ld, int bw, int64_t i, int64_t j) {
    return ab[i * ld + (j - i + bw)];
}

__global__ __launch_bounds__(256) void band_zero_kernel(double* __restrict__ ab, int64_t count) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        ab[t] = 0.0;
}

// One thread per block: its 9 values into the band.
__global__ __launch_bounds__(256) void band_fill_kernel(double* __restrict__ ab, int64_t ld, int bw,
                                                        int64_t num_blocks, const int* __restrict__ blk_row,
                                                        const int* __restrict__ colidx,
                                                        const double* __restrict__ S) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= num_blocks) return;
    const int64_t I = blk_row[k], J = colidx[k];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) at(ab, ld, bw, 3 * I + r, 3 * J + c) = S[9 * k + 3 * r + c];
}

// Panel [k0, k1) factored in place by one workgroup (global memory; every step ends with a barrier).
__global__ __launch_bounds__(1024) void band_panel_kernel(double* __restrict__ ab, int64_t n, int bw, int64_t ld,
                                                          int64_t k0, int64_t k1) {
    for (int64_t c = k0; c < k1; ++c) {
        // l(i, c) is non-zero for i in (c, c + bw], u(c, j) for j in (c, c + bw]; every product lands in the band
        const int64_t iend = min(n - 1, c + bw);  // last row with a multiplier in column c
        const int64_t jend = min(n - 1, c + bw);  // last column of row c in the band
        const double piv = atc(ab, ld, bw, c, c);
        for (int64_t i = c + 1 + threadIdx.x; i <= iend; i += blockDim.x) at(ab, ld, bw, i, c) /= piv;
        __syncthreads();
        // rank-1 update restricted to the panel's columns (c, min(k1 - 1, jend)] for rows (c, iend] ...
        const int64_t ncol = min(k1 - 1, jend) - c, nrow = iend - c;
        for (int64_t t = threadIdx.x; t < ncol * nrow; t += blockDim.x) {
            const int64_t i = c + 1 + t % nrow, j = c + 1 + t / nrow;
            at(ab, ld, bw, i, j) -= atc(ab, ld, bw, i, c) * atc(ab, ld, bw, c, j);
        }
        // ... and to the panel's rows (c, min(k1 - 1, iend)] for columns [k1, jend] (U12)
        const int64_t nrow2 = min(k1 - 1, iend) - c;
        const int64_t nu = jend >= k1 ? jend - k1 + 1 : 0;
        for (int64_t t = threadIdx.x; t < nrow2 * nu; t += blockDim.x) {
            const int64_t i = c + 1 + t / nu, j = k1 + t % nu;
            at(ab, ld, bw, i, j) -= atc(ab, ld, bw, i, c) * atc(ab, ld, bw, c, j);
        }
        __syncthreads();
    }
}

// Trailing update a(i, j) -= sum_c l(i, c) u(c, j) over the panel [k0, k1), i, j in [k1, k1 + m): one 64 x 64
// output tile per workgroup, 4 x 4 outputs per thread, panel factors staged in LDS.
__global__ __launch_bounds__(256) void band_update_kernel(double* __restrict__ ab, int bw, int64_t ld, int64_t k0,
                                                          int64_t k1, int64_t m) {
    __shared__ double Ls[kTile][kLuPanel + 1];
    __shared__ double Us[kLuPanel][kTile + 1];
    const int64_t i0 = k1 + (int64_t)blockIdx.y * kTile, j0 = k1 + (int64_t)blockIdx.x * kTile;
    const int64_t iend = k1 + m, jend = k1 + m;
    const int nb = (int)(k1 - k0);
    const int tid = threadIdx.x;
    for (int t = tid; t < kTile * kLuPanel; t += 256) {
        const int r = t / kLuPanel, c = t % kLuPanel;
        const int64_t i = i0 + r, col = k0 + c;
        double v = 0.0;
        if (c < nb && i < iend && i - col <= bw) v = atc(ab, ld, bw, i, col);
        Ls[r][c] = v;
    }
    for (int t = tid; t < kLuPanel * kTile; t += 256) {
        const int c = t / kTile, q = t % kTile;
        const int64_t row = k0 + c, j = j0 + q;
        double v = 0.0;
        if (c < nb && j < jend && j - row <= bw) v = atc(ab, ld, bw, row, j);
        Us[c][q] = v;
    }
    __syncthreads();
    const int tx = tid % 16, ty = tid / 16;
    double acc[4][4] = {};
    for (int c = 0; c < nb; ++c) {
        double l[4], u[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) l[a] = Ls[ty + 16 * a][c];
#pragma unroll
        for (int b = 0; b < 4; ++b) u[b] = Us[c][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t i = i0 + ty + 16 * a;
        if (i >= iend) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t j = j0 + tx + 16 * b;
            if (j >= jend) continue;
            at(ab, ld, bw, i, j) -= acc[a][b];
        }
    }
}

// x = U^-1 L^-1 b for column blockIdx.x, 1024 threads. Rows go in blocks of 64: 16 lanes per row sum the row's
// products with the entries solved before the block (a contiguous stretch of the band row, read coalesced, four
// independent accumulators per lane, then a butterfly over the 16 lanes), while the block's own triangle is
// staged in LDS; wavefront 0 then solves the 64 x 64 triangle from LDS with the solved entries broadcast lane to
// lane. Sums are in a fixed order.
constexpr int kSolveThreads = 1024;
__global__ __launch_bounds__(kSolveThreads) void band_solve_kernel(const double* __restrict__ ab, int64_t n, int bw,
                                                                   int64_t ld, const double* b, int64_t ldb, double* x,
                                                                   int64_t ldx) {
    __shared__ double T[kTile][kTile + 1];
    __shared__ double rs[kTile];
    const double* bc = b + (int64_t)blockIdx.x * ldb;
    double* xc = x + (int64_t)blockIdx.x * ldx;
    const int tid = threadIdx.x, row = tid >> 4, q = tid & 15, lane = tid & 63;
    auto row_sum = [&](double s) {
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (q == 0) rs[row] = s;
    };
    // forward: y = L^-1 b (unit diagonal), y written to xc
    for (int64_t r0 = 0; r0 < n; r0 += kTile) {
        for (int t = tid; t < kTile * kTile; t += kSolveThreads) {
            const int l = t >> 6, k = t & 63;
            const int64_t ii = r0 + l, kk = r0 + k;
            T[l][k] = (ii < n && kk < ii && ii - kk <= bw) ? atc(ab, ld, bw, ii, kk) : 0.0;
        }
        const int64_t i = r0 + row;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (i < n) {
            const int64_t kb = max((int64_t)0, i - bw);
            const double* arow = ab + i * ld + (kb - i + bw);  // L(i, kb), L(i, kb + 1), ...
            const double* xk = xc + kb;
            const int cnt = (int)(r0 - kb);  // solved entries k in [kb, r0)
            int k = q;
            for (; k + 48 < cnt; k += 64) {
                s0 += arow[k] * xk[k];
                s1 += arow[k + 16] * xk[k + 16];
                s2 += arow[k + 32] * xk[k + 32];
                s3 += arow[k + 48] * xk[k + 48];
            }
            for (; k < cnt; k += 16) s0 += arow[k] * xk[k];
        }
        row_sum((s0 + s1) + (s2 + s3));
        __syncthreads();
        if (tid < 64) {
            const int64_t ii = r0 + lane;
            double y = ii < n ? bc[ii] - rs[lane] : 0.0;
            for (int k = 0; k < kTile; ++k) {
                const double yk = __shfl(y, k, 64);
                if (lane > k) y -= T[lane][k] * yk;
            }
            if (ii < n) xc[ii] = y;
        }
        __syncthreads();
    }
    // backward: x = U^-1 y
    const int64_t nblk = (n + kTile - 1) / kTile;
    for (int64_t blk = nblk - 1; blk >= 0; --blk) {
        const int64_t r0 = blk * kTile, r1 = min(n, r0 + kTile);
        for (int t = tid; t < kTile * kTile; t += kSolveThreads) {
            const int l = t >> 6, k = t & 63;
            const int64_t ii = r0 + l, kk = r0 + k;
            T[l][k] = (ii < r1 && kk < r1 && kk >= ii && kk - ii <= bw) ? atc(ab, ld, bw, ii, kk) : 0.0;
        }
        const int64_t i = r0 + row;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (i < r1) {
            const int64_t ke = min(n - 1, i + bw);
            const double* arow = ab + i * ld + (r1 - i + bw);  // U(i, r1), U(i, r1 + 1), ...
            const double* xk = xc + r1;
            const int cnt = (int)(ke - r1 + 1);  // solved entries k in [r1, ke]
            int k = q;
            for (; k + 48 < cnt; k += 64) {
                s0 += arow[k] * xk[k];
                s1 += arow[k + 16] * xk[k + 16];
                s2 += arow[k + 32] * xk[k + 32];
                s3 += arow[k + 48] * xk[k + 48];
            }
            for (; k < cnt; k += 16) s0 += arow[k] * xk[k];
        }
        row_sum((s0 + s1) + (s2 + s3));
        __syncthreads();
        if (tid < 64) {
            const int64_t ii = r0 + lane;
            const int m = (int)(r1 - r0);
            double acc = ii < r1 ? xc[ii] - rs[lane] : 0.0;
            double xv = 0.0;
            for (int k = m - 1; k >= 0; --k) {
                if (lane == k) xv = acc / T[k][k];
                const double xkv = __shfl(xv, k, 64);
                if (lane < k) acc -= T[lane][k] * xkv;
            }
            if (ii < r1) xc[ii] = xv;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void identity_columns_kernel(int64_t n, int64_t first, int count,
                                                               double* __restrict__ out) {
    const int64_t total = n * (int64_t)count;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        out[t] = (i == first + j) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void zero_bc_rows_cols_kernel(int64_t n, int64_t first, int count,
                                                                const uint8_t* __restrict__ mask,
                                                                double* __restrict__ out) {
    const int64_t total = n * (int64_t)count;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        if (mask[i] || mask[first + j]) out[t] = 0.0;
    }
}

__global__ __launch_bounds__(256) void zero_masked_rows_kernel(int64_t n, int count, const uint8_t* __restrict__ mask,
                                                               double* __restrict__ x) {
    const int64_t total = n * (int64_t)count;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
        if (mask[t % n]) x[t] = 0.0;
}

unsigned grid_for(int64_t count, int per) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + per - 1) / per, 65536));
}

}  // namespace

size_t band_bytes(int64_t n, int bw) { return (size_t)n * (size_t)(2 * bw + 1) * sizeof(double); }

void launch_band_fill(DeviceBandLU& lu, int64_t num_blocks, const int* blk_row, const int* colidx, const double* S,
                      hipStream_t s) {
    const int64_t count = lu.n * lu.ld;
    launch_k(band_zero_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, lu.ab, count);
    if (num_blocks > 0)
        launch_k(band_fill_kernel, dim3((unsigned)((num_blocks + 255) / 256)), dim3(256), 0, s, lu.ab, lu.ld, lu.bw,
                 num_blocks, blk_row, colidx, S);
}

void launch_band_factor(DeviceBandLU& lu, hipStream_t s) {
    for (int64_t k0 = 0; k0 < lu.n; k0 += kLuPanel) {
        const int64_t k1 = std::min(lu.n, k0 + kLuPanel);
        launch_k(band_panel_kernel, dim3(1), dim3(1024), 0, s, lu.ab, lu.n, lu.bw, lu.ld, k0, k1);
        const int64_t m = std::min(lu.n, k1 + lu.bw) - k1;
        if (m > 0) {
            const unsigned tiles = (unsigned)((m + kTile - 1) / kTile);
            launch_k(band_update_kernel, dim3(tiles, tiles), dim3(256), 0, s, lu.ab, lu.bw, lu.ld, k0, k1, m);
        }
    }
}

void launch_band_solve(const DeviceBandLU& lu, int nrhs, const double* b, int64_t ldb, double* x, int64_t ldx,
                       hipStream_t s) {
    if (nrhs <= 0 || lu.n == 0) return;
    launch_k(band_solve_kernel, dim3((unsigned)nrhs), dim3(kSolveThreads), 0, s, (const double*)lu.ab, lu.n, lu.bw,
             lu.ld, b, ldb, x, ldx);
}

void launch_identity_columns(int64_t n, int64_t first, int count, double* out, hipStream_t s) {
    launch_k(identity_columns_kernel, dim3(grid_for(n * (int64_t)count, 256)), dim3(256), 0, s, n, first, count, out);
}

void launch_zero_bc_rows_cols(int64_t n, int64_t first, int count, const uint8_t* mask, double* out, hipStream_t s) {
    launch_k(zero_bc_rows_cols_kernel, dim3(grid_for(n * (int64_t)count, 256)), dim3(256), 0, s, n, first, count, mask,
             out);
}

// Partial products of a dense column-major C: workgroup (row tile, column chunk), one thread per row, the
// chunk's columns in order; the chunk's x values staged in LDS.
__global__ __launch_bounds__(256) void dense_gemv_part_kernel(int64_t n, const double* __restrict__ C,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ part) {
    __shared__ double xs[kGemvCols];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.y * kGemvCols;
    const int64_t j1 = min(n, j0 + kGemvCols);
    if (threadIdx.x < kGemvCols && j0 + threadIdx.x < j1) xs[threadIdx.x] = x[j0 + threadIdx.x];
    __syncthreads();
    if (i >= n) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const double* col = C + j0 * n + i;
    const int m = (int)(j1 - j0);
    int k = 0;
    for (; k + 3 < m; k += 4) {
        a0 += col[(int64_t)k * n] * xs[k];
        a1 += col[(int64_t)(k + 1) * n] * xs[k + 1];
        a2 += col[(int64_t)(k + 2) * n] * xs[k + 2];
        a3 += col[(int64_t)(k + 3) * n] * xs[k + 3];
    }
    for (; k < m; ++k) a0 += col[(int64_t)k * n] * xs[k];
    part[(int64_t)blockIdx.y * n + i] = (a0 + a1) + (a2 + a3);
}

__global__ __launch_bounds__(256) void dense_gemv_sum_kernel(int64_t n, int64_t chunks,
                                                             const double* __restrict__ part, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int64_t c = 0; c < chunks; ++c) acc += part[c * n + i];
    y[i] = acc;
}

void launch_dense_gemv(int64_t n, const double* C, const double* x, double* part, double* y, hipStream_t s) {
    if (n <= 0) return;
    const int64_t chunks = gemv_chunks(n);
    launch_k(dense_gemv_part_kernel, dim3(grid_for(n, 256), (unsigned)chunks), dim3(256), 0, s, n, C, x, part);
    launch_k(dense_gemv_sum_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, chunks, (const double*)part, y);
}

void launch_zero_masked_rows(int64_t n, int count, const uint8_t* mask, double* x, hipStream_t s) {
    launch_k(zero_masked_rows_kernel, dim3(grid_for(n * (int64_t)count, 256)), dim3(256), 0, s, n, count, mask, x);
}

}  // namespace femgpu
