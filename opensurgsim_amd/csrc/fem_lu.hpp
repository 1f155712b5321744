// Direct solver of the assembled operator: banded factorization without pivoting in FP64, the GPU form of
// LinearSparseSolveAndInverseLU (Math/LinearSparseSolveAndInverse.cpp:31-97, Eigen's SparseLU there). The
// operator of this path (M/dt + D + dt K with boundary rows and columns replaced by the identity, the lower
// triangle mirrored) is symmetric with a positive definite symmetric part, for which elimination without
// pivoting is stable; it is eliminated as LDL^T and stored as the LU factors (U = D L^T). The node numbering of a
// block mesh keeps the bandwidth at 3 (max |I - J| over coupled nodes) + 2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femgpu {

// Band storage, row-major: entry (i, j) with |i - j| <= bw at ab[i * ld + (j - i + bw)], ld = 2 bw + 1. After
// launch_band_factor it holds the unit lower factor L (below the diagonal) and U (on and above it), and inv the
// inverses of the 64 x 64 diagonal blocks of L and U (block b: L's at inv + 2 b 64^2, U's at inv + (2 b + 1)
// 64^2, row-major, rows past n padded with the identity). inv lies in the same allocation, after the band.
struct DeviceBandLU {
    int64_t n = 0;
    int bw = 0;
    int64_t ld = 0;
    double* ab = nullptr;
    double* inv = nullptr;
    size_t capacity = 0;  // doubles allocated
};

// Number of 64-row blocks, and the bytes the band storage plus the block inverses of an n x n matrix of
// bandwidth bw need.
int64_t band_blocks(int64_t n);
size_t band_bytes(int64_t n, int bw);

// ab = the 3x3 block CSR operator S (9 doubles per block, row-major in the block; block k in block row
// blk_row[k], block column colidx[k]) scattered into band storage of bandwidth bw (zero elsewhere).
void launch_band_fill(DeviceBandLU& lu, int64_t num_blocks, const int* blk_row, const int* colidx, const double* S,
                      hipStream_t s);
// In-place factorization of the band storage (the lower triangle is read): panels of 32 columns, each the
// panel's rows eliminated against its diagonal block and the rank-32 update of the trailing bw x bw lower
// triangle by LDS-tiled workgroups; then the inverses of the diagonal blocks.
void launch_band_factor(DeviceBandLU& lu, hipStream_t s);
// x = A^-1 b for nrhs columns (column j at b + j * ldb and x + j * ldx; x may alias b): forward substitution
// with L and back substitution with U by 64-row blocks, one workgroup per 16 columns (4 below 64 columns).
// identity_first >= 0: b holds columns identity_first .. of the identity, and the forward pass skips the rows
// above each workgroup's first column (their solutions are zero).
void launch_band_solve(const DeviceBandLU& lu, int nrhs, const double* b, int64_t ldb, double* x, int64_t ldx,
                       hipStream_t s, int64_t identity_first = -1);
// Columns [first, first + count) of the identity (n rows each, column-major with leading dimension n).
void launch_identity_columns(int64_t n, int64_t first, int count, double* out, hipStream_t s);
// Rows and columns of the boundary DOFs (mask != 0) of an n x count column block (columns first.. of an n x n
// matrix) set to zero: OdeState::applyBoundaryConditionsToMatrix(&C, false).
void launch_zero_bc_rows_cols(int64_t n, int64_t first, int count, const uint8_t* mask, double* out, hipStream_t s);

// Rows of the boundary DOFs (mask != 0) of count columns of n values (column-major) set to zero.
void launch_zero_masked_rows(int64_t n, int count, const uint8_t* mask, double* x, hipStream_t s);

// y = C x for a dense n x n column-major C: partial sums over column chunks of kGemvCols (part holds
// gemv_chunks(n) x n doubles), then each row's partials added in chunk order (deterministic).
constexpr int kGemvCols = 128;
inline int64_t gemv_chunks(int64_t n) { return (n + kGemvCols - 1) / kGemvCols; }
void launch_dense_gemv(int64_t n, const double* C, const double* x, double* part, double* y, hipStream_t s);

}  // namespace femgpu
