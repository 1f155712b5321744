// Banded factorization without pivoting in FP64 (see fem_lu.hpp).
//
// The operator is symmetric (its lower triangle mirrored), so the elimination runs on the lower triangle only
// (LDL^T) and writes the LU form the solves read: L (unit lower) below the diagonal, U = D L^T on and above it.
// Row c of U is column c of the eliminated lower triangle, the value the elimination divides by the pivot.
// Right-looking, by panels of kLuPanel columns. For panel [k0, k1):
//   1. its diagonal block is eliminated in LDS (for the first panel by its own launch, for the others by the first
//      workgroup of the previous panel's trailing update, which updates that block itself);
//   2. the rows below, k1 .. k1 + bw - 1, are eliminated against it, one lane per row: l(i, c) and u(c, i);
//   3. the trailing block a(i, j), i >= j in [k1, k1 + bw), takes -= L21 U12 by 64 x 64 output tiles staged
//      through LDS (lower tiles only: the upper band is rewritten from L by step 2 before anything reads it).
// Each entry takes the column-by-column updates in the same order as the column form of the restated reference
// solver, with the panel's updates of the trailing block summed before they are applied: results agree with
// unblocked LU to rounding.
//
// Substitution: the 64 x 64 diagonal blocks of L and U are inverted once per factorization. A solve walks the
// rows in blocks of 64: the block's products with the entries solved before it (up to bw per row) as an LDS-tiled
// product over 64-entry tiles (the next tile prefetched into registers), then the block's triangle as a product
// with its inverse. One workgroup per 16 (or 4) right-hand sides; sums in a fixed order.
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

// Elimination of an nb x nb diagonal block held in LDS (T: kLuPanel rows of kLuPanel + 1 doubles, its lower
// triangle) by the whole workgroup: step c subtracts (a(i, c) / a(c, c)) u(c, j) from the lower entries (i, j),
// c < j <= i, where u(c, j) = T[j][c] is the eliminated column value (the division as a product with the pivot's
// reciprocal). On return T's lower triangle holds the eliminated values: u(c, i) = T[i][c], pivots on the diagonal.
__device__ void eliminate_diag_block(int nb, double (*T)[kLuPanel + 1]) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int c = 0; c < nb; ++c) {
        const int w = nb - 1 - c;  // entries below the pivot
        const double rp = 1.0 / T[c][c];
        for (int e = tid; e < w * w; e += nt) {
            const int i = c + 1 + e / w, j = c + 1 + e % w;
            if (j <= i) T[i][j] -= (T[i][c] * rp) * T[j][c];
        }
        __syncthreads();
    }
}

// The eliminated block into the band at k0: L (l(i, c) = u(c, i) / u(c, c), as a product with the reciprocal)
// below the diagonal, U on and above it.
__device__ void write_diag_block(double* __restrict__ ab, int bw, int64_t ld, int64_t k0, int nb,
                                 const double (*T)[kLuPanel + 1]) {
    for (int e = threadIdx.x; e < kLuPanel * kLuPanel; e += blockDim.x) {
        const int i = e / kLuPanel, c = e % kLuPanel;
        if (i >= nb || c > i || i - c > bw) continue;
        if (c == i) {
            at(ab, ld, bw, k0 + i, k0 + i) = T[i][i];
        } else {
            at(ab, ld, bw, k0 + i, k0 + c) = T[i][c] * (1.0 / T[c][c]);
            at(ab, ld, bw, k0 + c, k0 + i) = T[i][c];
        }
    }
}

// The first panel's diagonal block: staged from the band, eliminated, written back.
__global__ __launch_bounds__(256) void ldl_diag_kernel(double* __restrict__ ab, int bw, int64_t ld, int64_t k0,
                                                       int nb) {
    __shared__ double T[kLuPanel][kLuPanel + 1];
    for (int e = threadIdx.x; e < kLuPanel * kLuPanel; e += blockDim.x) {
        const int i = e / kLuPanel, j = e % kLuPanel;
        T[i][j] = (i < nb && j <= i && i - j <= bw) ? atc(ab, ld, bw, k0 + i, k0 + j) : 0.0;
    }
    __syncthreads();
    eliminate_diag_block(nb, T);
    write_diag_block(ab, bw, ld, k0, nb, T);
}

// Rows k1 + 64 blockIdx.x + lane (< k1 + m) of panel [k0, k1 = k0 + nb), eliminated against the panel's factored
// diagonal block (U11 from the upper band, staged in LDS with the pivots' reciprocals), column by column:
// s_k = the row's entry k with the updates of columns < k applied, l(i, k) = s_k / u(k, k) (as a product with the
// reciprocal) into row i's lower band and s_k = u(k, i) into row k's upper band, then l(i, k) u(k, c) subtracted
// from the entries c > k (independent of each other).
__global__ __launch_bounds__(64) void ldl_panel_kernel(double* __restrict__ ab, int bw, int64_t ld, int64_t k0,
                                                       int nb, int64_t m) {
    __shared__ double Us[kLuPanel][kLuPanel + 1];
    __shared__ double rp[kLuPanel];
    const int lane = threadIdx.x;
    for (int e = lane; e < kLuPanel * kLuPanel; e += 64) {
        const int c = e / kLuPanel, j = e % kLuPanel;
        Us[c][j] = (c < nb && j < nb && j >= c && j - c <= bw) ? atc(ab, ld, bw, k0 + c, k0 + j) : 0.0;
    }
    if (lane < kLuPanel) rp[lane] = lane < nb ? 1.0 / atc(ab, ld, bw, k0 + lane, k0 + lane) : 0.0;
    __syncthreads();
    const int64_t k1 = k0 + nb;
    const int64_t i = k1 + (int64_t)blockIdx.x * 64 + lane;
    if (i >= k1 + m) return;
    double r[kLuPanel];
#pragma unroll
    for (int c = 0; c < kLuPanel; ++c) {
        const bool in = c < nb && i - (k0 + c) <= bw;
        const double v = ab[in ? i * ld + (k0 + c - i + bw) : 0];
        r[c] = in ? v : 0.0;
    }
    // Row k of U11 and the pivot's reciprocal come from LDS one step ahead (cur: step k, nxt: step k + 1); the
    // scheduling barrier keeps those loads ahead of step k's products.
    double ua[kLuPanel + 1], ub[kLuPanel + 1];  // [c]: u(k, c) for c > k; [kLuPanel]: 1 / u(k, k)
    auto fetch = [&](int k, double (&u)[kLuPanel + 1]) {
#pragma unroll
        for (int c = k + 1; c < kLuPanel; ++c) u[c] = Us[k][c];
        u[kLuPanel] = rp[k];
    };
    auto step = [&](int k, double (&cur)[kLuPanel + 1], double (&nxt)[kLuPanel + 1]) {
        if (k >= nb) return;
        if (k + 1 < kLuPanel) fetch(k + 1, nxt);
        __builtin_amdgcn_sched_barrier(0);
        const double sk = r[k];
        const double l = sk * cur[kLuPanel];
        if (i - (k0 + k) <= bw) {  // beyond the band s and l are zero
            at(ab, ld, bw, k0 + k, i) = sk;
            at(ab, ld, bw, i, k0 + k) = l;
        }
#pragma unroll
        for (int c = k + 1; c < kLuPanel; ++c) r[c] -= l * cur[c];
    };
    fetch(0, ua);
#pragma unroll
    for (int k = 0; k < kLuPanel; k += 2) {
        step(k, ua, ub);
        step(k + 1, ub, ua);
    }
}

// Trailing update a(i, j) -= sum_c l(i, c) u(c, j) over the panel [k0, k1), i >= j in [k1, k1 + m). Workgroup 0
// takes the next panel's diagonal block [k1, k1 + next_nb)^2: its update (the same sums in the same order), then
// its elimination and its L and U into the band. Workgroups 1.. take the 64 x 64 lower output tiles ((ti, tj),
// tj <= ti, row by row; tile (0, 0) leaves the diagonal block to workgroup 0), 4 x 4 outputs per thread, the
// panel's L21 and U12 staged in LDS.
__global__ __launch_bounds__(256) void ldl_update_kernel(double* __restrict__ ab, int bw, int64_t ld, int64_t k0,
                                                         int64_t k1, int64_t m, int next_nb) {
    __shared__ double Ls[kTile][kLuPanel + 1];
    __shared__ double Us[kLuPanel][kTile + 1];
    const int nb = (int)(k1 - k0);
    const int tid = threadIdx.x;
    const int64_t iend = k1 + m, jend = k1 + m;
    if (blockIdx.x == 0) {
        if (next_nb <= 0) return;
        double (*T)[kLuPanel + 1] = &Ls[kLuPanel];  // rows 32..63 of Ls
        for (int t = tid; t < kLuPanel * kLuPanel; t += 256) {
            const int r = t / kLuPanel, c = t % kLuPanel;
            const int64_t i = k1 + r, col = k0 + c;
            Ls[r][c] = (c < nb && r < next_nb && i < iend && i - col <= bw) ? atc(ab, ld, bw, i, col) : 0.0;
            const int64_t row = k0 + r, j = k1 + c;  // U12 entry (r, c) of the block's columns
            Us[r][c] = (r < nb && c < next_nb && j < jend && j - row <= bw) ? atc(ab, ld, bw, row, j) : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < kLuPanel * kLuPanel; e += 256) {
            const int i = e / kLuPanel, j = e % kLuPanel;
            const bool in = i < next_nb && j <= i && i - j <= bw;
            const double old = ab[in ? (k1 + i) * ld + (j - i + bw) : 0];
            double acc = 0.0;  // Ls and Us are zero past nb: the sum equals the tiles' sum over c < nb
#pragma unroll 1
            for (int c0 = 0; c0 < kLuPanel; c0 += 8) {
                double l[8], u[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    l[q] = Ls[i][c0 + q];
                    u[q] = Us[c0 + q][j];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) acc += l[q] * u[q];
            }
            T[i][j] = in ? old - acc : 0.0;
        }
        __syncthreads();
        eliminate_diag_block(next_nb, T);
        write_diag_block(ab, bw, ld, k1, next_nb, T);
        return;
    }
    const int b = (int)blockIdx.x - 1;
    int ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    const int tj = b - ti * (ti + 1) / 2;
    const int64_t i0 = k1 + (int64_t)ti * kTile, j0 = k1 + (int64_t)tj * kTile;
    const int tx = tid % 16, ty = tid / 16;
    // entries (i, j) this tile writes: in the band, inside the trailing block, not in the next diagonal block
    auto mine = [&](int64_t i, int64_t j) {
        return i < iend && j < jend && j - i <= bw && i - j <= bw && !(i < k1 + next_nb && j < k1 + next_nb);
    };
    double old[4][4];  // the outputs, fetched first
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int64_t i = i0 + ty + 16 * a, j = j0 + tx + 16 * bb;
            const bool in = mine(i, j);
            const double v = ab[in ? i * ld + (j - i + bw) : 0];
            old[a][bb] = in ? v : 0.0;
        }
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
    double acc[4][4] = {};  // Ls and Us are zero past nb, so the loop runs over every staged column
#pragma unroll 1
    for (int c0 = 0; c0 < kLuPanel; c0 += 4) {
        double l[4][4], u[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int a = 0; a < 4; ++a) l[q][a] = Ls[ty + 16 * a][c0 + q];
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) u[q][bb] = Us[c0 + q][tx + 16 * bb];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) acc[a][bb] += l[q][a] * u[q][bb];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int64_t i = i0 + ty + 16 * a, j = j0 + tx + 16 * bb;
            if (mine(i, j)) at(ab, ld, bw, i, j) = old[a][bb] - acc[a][bb];
        }
}

// Inverse of the 64 x 64 diagonal block blockIdx.x of L (blockIdx.y = 0, unit lower) or U (1, upper), row-major
// at inv + (2 blockIdx.x + blockIdx.y) 64^2; rows past n padded with the identity. Lane c forms column c by
// substitution on the identity column.
__global__ __launch_bounds__(64) void band_block_inverse_kernel(const double* __restrict__ ab, int64_t n, int bw,
                                                                int64_t ld, double* __restrict__ inv) {
    __shared__ double T[kTile][kTile];  // read at one address per step by every lane
    __shared__ double X[kTile][kTile];  // lane c's column
    const int64_t r0 = (int64_t)blockIdx.x * kTile;
    const int mb = (int)min((int64_t)kTile, n - r0);
    const bool upper = blockIdx.y == 1;
    const int c = threadIdx.x;
    for (int e = c; e < kTile * kTile; e += 64) {
        const int r = e >> 6, k = e & 63;
        double v = (r == k) ? 1.0 : 0.0;
        if (r < mb && k < mb && (upper ? (k >= r && k - r <= bw) : (k < r && r - k <= bw)))
            v = atc(ab, ld, bw, r0 + r, r0 + k);
        T[r][k] = v;
    }
    __syncthreads();
    if (!upper) {
        for (int r = 0; r < kTile; ++r) {
            double s = (r == c) ? 1.0 : 0.0;
            for (int k = c; k < r; ++k) s -= T[r][k] * X[k][c];
            X[r][c] = s;
        }
    } else {
        for (int r = kTile - 1; r >= 0; --r) {
            double s = (r == c) ? 1.0 : 0.0;
            for (int k = r + 1; k <= c; ++k) s -= T[r][k] * X[k][c];
            X[r][c] = s / T[r][r];
        }
    }
    __syncthreads();
    double* out = inv + ((int64_t)blockIdx.x * 2 + blockIdx.y) * (kTile * kTile);
    for (int e = c; e < kTile * kTile; e += 64) out[e] = X[e >> 6][e & 63];
}

// x = U^-1 L^-1 b for columns [R blockIdx.x, R blockIdx.x + R) of nrhs, 256 threads. Products over a 64-entry
// tile are split by entries: wavefront g takes entries [16 g, 16 g + 16) of every row and column, and the four
// partial sums of a row block are added in wavefront order; thread (row l = tid % 64, g) then owns columns
// [g R / 4, (g + 1) R / 4) of the row. identity_first >= 0: b holds columns identity_first + j of the identity,
// whose forward solutions are zero above row identity_first + j; the forward pass starts at the row block of the
// first column (rows above it written zero). x may alias b.
template <int R>
__global__ __launch_bounds__(256) void band_solve_blocked_kernel(const double* __restrict__ ab,
                                                                 const double* __restrict__ inv, int64_t n, int bw,
                                                                 int64_t ld, int nrhs, const double* b, int64_t ldb,
                                                                 double* x, int64_t ldx, int64_t identity_first) {
    constexpr int CPT = R / 4;               // columns each thread finishes
    constexpr int PA = kTile * kTile / 256;  // tile values each thread stages
    constexpr int PY = kTile * R / 256;
    constexpr int KW = kTile / 4;            // tile entries per wavefront
    static_assert(4 * R * kTile <= kTile * (kTile + 1), "partial sums fit in At");
    __shared__ double At[kTile][kTile + 1];  // At[k][row]; also the wavefronts' partial sums
    __shared__ double Ys[kTile][R + 2];      // Ys[k][column] (rows 16-byte aligned)
    __shared__ double Rs[kTile][R + 2];      // Rs[row][column]
    const int tid = threadIdx.x, l = tid & 63, g = tid >> 6, kk0 = tid & 63;
    const int64_t j0 = (int64_t)blockIdx.x * R;
    const int64_t nblk = (n + kTile - 1) / kTile;
    int64_t sb = 0;
    if (identity_first >= 0) {
        sb = (identity_first + j0) / kTile;
        const int64_t rows = sb * kTile;
        for (int64_t e = tid; e < rows * R; e += 256) {
            const int64_t cc = e / rows, i = e - cc * rows;
            if (j0 + cc < nrhs) x[(j0 + cc) * ldx + i] = 0.0;
        }
    }
    double pa[PA], py[PY];
    // Tile [kt, kt + kn) of row block r0 into registers (thread: entry kk0, rows g + 4 q; columns g + 4 q of x),
    // loads issued unconditionally from clamped addresses. The forward pass reads L (k < i), the backward U (k > i).
    auto load_tile = [&](int64_t r0, int64_t rend, int64_t kt, int kn, bool lower) {
        const int64_t k = kt + kk0;
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int64_t i = r0 + g + 4 * q;
            const bool in = kk0 < kn && i < rend && (lower ? i - k <= bw : k - i <= bw);
            const double v = ab[in ? i * ld + (k - i + bw) : 0];
            pa[q] = in ? v : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            const int64_t j = j0 + g + 4 * q;
            const bool in = kk0 < kn && j < nrhs;
            const double v = x[in ? j * ldx + k : 0];
            py[q] = in ? v : 0.0;
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int q = 0; q < PA; ++q) At[kk0][g + 4 * q] = pa[q];
#pragma unroll
        for (int q = 0; q < PY; ++q) Ys[kk0][g + 4 * q] = py[q];
    };
    // acc[c] += sum over this wavefront's entries of At[k][l] S[k][c]
    auto kpart = [&](const double (*S)[R + 2], double* acc) {
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int kk = g * KW + u;
            const double a = At[kk][l];
#pragma unroll
            for (int c = 0; c < R; ++c) acc[c] += a * S[kk][c];
        }
    };
    // out[q] = the four wavefronts' partials of column g CPT + q of row l, added in wavefront order. At must be
    // free on entry; it is free again on return.
    auto reduce = [&](const double* acc, double* out) {
        double* P = &At[0][0];
#pragma unroll
        for (int c = 0; c < R; ++c) P[(g * R + c) * kTile + l] = acc[c];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int c = g * CPT + q;
            out[q] = ((P[c * kTile + l] + P[(R + c) * kTile + l]) + P[(2 * R + c) * kTile + l]) +
                     P[(3 * R + c) * kTile + l];
        }
        __syncthreads();
    };
    // tot = the row block's band products with the solved entries [kbeg, kend); At free on return
    auto products = [&](int64_t r0, int64_t rend, int64_t kbeg, int64_t kend, bool lower, double* tot) {
        double acc[R];
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] = 0.0;
        if (kbeg < kend) {
            load_tile(r0, rend, kbeg, (int)min((int64_t)kTile, kend - kbeg), lower);
            for (int64_t kt = kbeg; kt < kend; kt += kTile) {
                __syncthreads();  // the last tile's readers are done
                store_tile();
                __syncthreads();
                const int64_t kn = kt + kTile;
                if (kn < kend) load_tile(r0, rend, kn, (int)min((int64_t)kTile, kend - kn), lower);
                kpart(Ys, acc);
            }
            __syncthreads();
        }
        reduce(acc, tot);
    };
    // x rows [r0, rend) of the block's columns = inverse block (row-major at blk) times Rs
    auto triangle = [&](const double* __restrict__ blk, int64_t r0, int64_t rend) {
#pragma unroll
        for (int q = 0; q < PA; ++q) At[kk0][g + 4 * q] = blk[(g + 4 * q) * kTile + kk0];
        __syncthreads();  // At and Rs complete
        double acc[R];
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] = 0.0;
        kpart(Rs, acc);
        __syncthreads();
        double y[CPT];
        reduce(acc, y);
        const int64_t i = r0 + l;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int64_t j = j0 + g * CPT + q;
            if (i < rend && j < nrhs) x[j * ldx + i] = y[q];
        }
        __syncthreads();  // x rows written before the next block reads them
    };
    double tot[CPT];
    for (int64_t rb = sb; rb < nblk; ++rb) {  // forward: y = L^-1 b
        const int64_t r0 = rb * kTile;
        products(r0, n, std::max(sb * kTile, r0 - bw), r0, true, tot);
        const int64_t i = r0 + l;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int64_t j = j0 + g * CPT + q;
            Rs[l][g * CPT + q] = (i < n && j < nrhs) ? b[j * ldb + i] - tot[q] : 0.0;
        }
        triangle(inv + rb * 2 * (kTile * kTile), r0, n);
    }
    for (int64_t rb = nblk - 1; rb >= 0; --rb) {  // backward: x = U^-1 y
        const int64_t r0 = rb * kTile, r1 = std::min(n, r0 + kTile);
        products(r0, r1, r1, std::min(n, r1 + bw), false, tot);
        const int64_t i = r0 + l;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int64_t j = j0 + g * CPT + q;
            Rs[l][g * CPT + q] = (i < r1 && j < nrhs) ? x[j * ldx + i] - tot[q] : 0.0;
        }
        triangle(inv + (rb * 2 + 1) * (kTile * kTile), r0, r1);
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

int64_t band_blocks(int64_t n) { return (n + kTile - 1) / kTile; }

size_t band_bytes(int64_t n, int bw) {
    return ((size_t)n * (size_t)(2 * bw + 1) + (size_t)band_blocks(n) * 2 * kTile * kTile) * sizeof(double);
}

void launch_band_fill(DeviceBandLU& lu, int64_t num_blocks, const int* blk_row, const int* colidx, const double* S,
                      hipStream_t s) {
    const int64_t count = lu.n * lu.ld;
    launch_k(band_zero_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, lu.ab, count);
    if (num_blocks > 0)
        launch_k(band_fill_kernel, dim3((unsigned)((num_blocks + 255) / 256)), dim3(256), 0, s, lu.ab, lu.ld, lu.bw,
                 num_blocks, blk_row, colidx, S);
}

void launch_band_factor(DeviceBandLU& lu, hipStream_t s) {
    if (lu.n <= 0) return;
    if (!lu.inv) throw std::runtime_error("band LU without block-inverse storage");
    launch_k(ldl_diag_kernel, dim3(1), dim3(256), 0, s, lu.ab, lu.bw, lu.ld, (int64_t)0,
             (int)std::min<int64_t>(lu.n, kLuPanel));
    for (int64_t k0 = 0; k0 < lu.n; k0 += kLuPanel) {
        const int64_t k1 = std::min(lu.n, k0 + kLuPanel);
        const int64_t m = std::min(lu.n, k1 + lu.bw) - k1;
        if (m <= 0) break;  // the last panel: its diagonal block is all there is
        launch_k(ldl_panel_kernel, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, s, lu.ab, lu.bw, lu.ld, k0,
                 (int)(k1 - k0), m);
        const int64_t tiles = (m + kTile - 1) / kTile;
        launch_k(ldl_update_kernel, dim3((unsigned)(1 + tiles * (tiles + 1) / 2)), dim3(256), 0, s, lu.ab, lu.bw, lu.ld,
                 k0, k1, m, (int)(std::min(lu.n, k1 + kLuPanel) - k1));
    }
    launch_k(band_block_inverse_kernel, dim3((unsigned)band_blocks(lu.n), 2), dim3(64), 0, s, (const double*)lu.ab,
             lu.n, lu.bw, lu.ld, lu.inv);
}

void launch_band_solve(const DeviceBandLU& lu, int nrhs, const double* b, int64_t ldb, double* x, int64_t ldx,
                       hipStream_t s, int64_t identity_first) {
    if (nrhs <= 0 || lu.n == 0) return;
    if (nrhs >= 64)
        launch_k(band_solve_blocked_kernel<16>, dim3((unsigned)((nrhs + 15) / 16)), dim3(256), 0, s,
                 (const double*)lu.ab, (const double*)lu.inv, lu.n, lu.bw, lu.ld, nrhs, b, ldb, x, ldx,
                 identity_first);
    else
        launch_k(band_solve_blocked_kernel<4>, dim3((unsigned)((nrhs + 3) / 4)), dim3(256), 0, s,
                 (const double*)lu.ab, (const double*)lu.inv, lu.n, lu.bw, lu.ld, nrhs, b, ldb, x, ldx,
                 identity_first);
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

// y[i] = the chunks' partials of row i added in chunk order; 64 rows per workgroup, loads 8 chunks at a time.
__global__ __launch_bounds__(64) void dense_gemv_sum_kernel(int64_t n, int64_t chunks,
                                                            const double* __restrict__ part, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    int64_t c = 0;
    for (; c + 8 <= chunks; c += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = part[(c + q) * n + i];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += v[q];
    }
    for (; c < chunks; ++c) acc += part[c * n + i];
    y[i] = acc;
}

void launch_dense_gemv(int64_t n, const double* C, const double* x, double* part, double* y, hipStream_t s) {
    if (n <= 0) return;
    if (n > (int64_t)65536 * 64) throw std::runtime_error("dense GEMV: more rows than one grid covers");
    const int64_t chunks = gemv_chunks(n);
    launch_k(dense_gemv_part_kernel, dim3(grid_for(n, 256), (unsigned)chunks), dim3(256), 0, s, n, C, x, part);
    launch_k(dense_gemv_sum_kernel, dim3(grid_for(n, 64)), dim3(64), 0, s, n, chunks, (const double*)part, y);
}

void launch_zero_masked_rows(int64_t n, int count, const uint8_t* mask, double* x, hipStream_t s) {
    launch_k(zero_masked_rows_kernel, dim3(grid_for(n * (int64_t)count, 256)), dim3(256), 0, s, n, count, mask, x);
}

}  // namespace femgpu
