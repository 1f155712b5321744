// Multi-column FP64 conjugate gradients on the assembled block operator: up to 8 right-hand sides solved
// together, each with its own Eigen ConjugateGradient recurrence, scalars and stopping test, reading the
// operator once per iteration for all of them. Used for compliance solves (C b column by column).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace femgpu {

// Per-column state. fin: the iteration in which the column finished (converged or capped); kMcgRunning
// while it iterates, -1 when it finished before the first iteration (zero right-hand side or already below
// the threshold). A column takes part in iteration k's kernels while fin >= k.
constexpr int kMcgRunning = 0x7fffffff;
struct McgColumn {
    double rz[2];      // r . z by iteration parity
    double bb;         // |b|^2
    double threshold;  // max(tol^2 |b|^2, realmin)
    double rr;         // |r|^2 at the last test
    int fin;
    int iters;         // Eigen's iteration count at exit
    int converged;
    int pad;
};

struct McgArgs {
    int kw;                 // columns in the batch: 1, 2, 4 or 8 (vectors interleaved, dof-major)
    int num_nodes;          // block rows
    const int* rowptr;      // block CSR of the operator
    const int* colidx;
    const int* diagblk;     // diagonal block of each row (-1: none)
    const double* S;        // 9 FP64 per block, row-major
    const double* b;        // [n][kw] right-hand sides (boundary rows already zero)
    double* x;              // [n][kw] solutions
    double *r, *z, *p, *q;  // [n][kw] work vectors
    double* dinv;           // [n] Jacobi preconditioner (1 / diagonal, 1 where it is 0)
    double* partial;        // [3][kCgBlocks][kw]
    McgColumn* col;         // [kw]
    double tol;
    int max_iter;
};

// Fixed device workspace of the allocation-free compliance entry point: a file-scope device array of
// kNoAllocScratchBytes per GPU, held from the library's first use of the GPU until the process ends.
// noalloc_scratch_address returns its address on `device` (which must be the current device), taken once per
// device with hipGetSymbolAddress and cached; it throws std::runtime_error when the lookup fails. Every call
// on one GPU shares the one array, so calls that overlap on that GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(96) << 20;
void* noalloc_scratch_address(int device);

// Byte offsets of the multi-column CG buffers in one workspace for n DOFs (8 columns); total bytes needed.
struct McgLayout {
    size_t b, x, r, z, p, q, dinv, partial, col, total;
};
McgLayout mcg_layout(size_t n);

void launch_mcg_init(const McgArgs& a, hipStream_t s);
// Column-major right-hand sides b (cnt columns of n) into the interleaved [n][kw] layout, boundary rows and
// padding columns zero; and back: out (cnt columns of n) from x [n][kw], boundary rows zero.
void launch_mcg_gather(int64_t n, int kw, int cnt, const double* b, const uint8_t* mask, double* dst, hipStream_t s);
void launch_mcg_scatter(int64_t n, int kw, int cnt, const double* x, const uint8_t* mask, double* out, hipStream_t s);
// dst = src with the boundary rows zero.
void launch_masked_copy(int64_t n, const double* src, const uint8_t* mask, double* dst, hipStream_t s);
// dst[idx[k]] = val[k] and dst[k] = src[idx[k]] for k < cnt (indices checked by the caller).
void launch_scatter_values(int cnt, const int64_t* idx, const double* val, double* dst, hipStream_t s);
void launch_gather_values(int cnt, const int64_t* idx, const double* src, double* dst, hipStream_t s);
void launch_mcg_iteration(const McgArgs& a, int k, hipStream_t s);

}  // namespace femgpu
