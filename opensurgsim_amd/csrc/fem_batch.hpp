// Batched conjugate gradient over many independent instances (disconnected meshes) in one set of
// launches: every instance keeps its own Eigen-CG scalars, stopping test and iteration count.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fem_kernels.hpp"

namespace femgpu {

constexpr int kBatchThreads = 256;  // one workgroup per chunk

// Rows (block rows = nodes) are cut into chunks that never straddle two instances; chunk c covers block
// rows [row0[c], row1[c]) of instance inst[c]; the chunks of instance i are inst_chunk[i] .. inst_chunk[i+1].
// Per-chunk partial sums live at partial[slot * num_chunks + c].
struct BatchTables {
    int num_instances;
    int num_chunks;
    const int* chunk_inst;
    const int* chunk_row0;
    const int* chunk_row1;
    const int* inst_chunk;   // [M + 1]
    const int* inst_maxit;   // [M] iteration cap per instance
    const int* inst_active;  // [M] 0 once the instance's state became non-finite (deactivated)
    // [M] 1: this solve leaves the instance alone (its x untouched, its scalars done); null: none. A batch in
    // the FP32 + refinement mode solves the instances flagged for FP64 in a second, FP64 pass.
    const int* inst_skip = nullptr;
    const int* live = nullptr;  // iteration kernels: workgroup b works on chunk live[b] (null: chunk b)
    int grid = 0;               // iteration kernels: workgroups launched (0: num_chunks)
};

// Device-side iteration counters, one pair per instance: the operator kernel reads kA and writes kV,
// the vector kernel reads kV and writes kA = kV + 1 (no kernel reads what it writes), so a captured
// sequence of iterations can be replayed.
struct BatchCg {
    BatchTables t;
    Operator op;
    bool fp32;
    const void* S;     // operator in the CG precision
    void* x;
    void* r;
    void* z;
    void* p[2];
    void* q;
    void* dinv;        // FP64: 1 / diagonal per DOF; FP32: 3x3 block-diagonal inverse, 9 per node
    double* partial;   // [4 * num_chunks]
    CgScalars* sc;     // [M] inner
    int* kA;           // [M]
    int* kV;           // [M]
};

// x = 0, r = b, dinv, z, per-instance |b|^2 and r.z; instances whose skip[i].done is set are left alone.
void launch_batch_cg_init(const BatchCg& a, const double* b, const CgScalars* skip, double tol, int lanes,
                          hipStream_t s);
// One CG iteration (operator kernel + vector kernel) for every instance not done.
void launch_batch_cg_iteration(const BatchCg& a, int lanes, hipStream_t s);
// FP64 refinement: r64 = b - S64 dv per instance (first: r64 = b), outer test per instance (marks
// inner done when passed), and dv += x (FP32) for instances still refining.
void launch_batch_residual(const BatchCg& a, const double* S64, const double* b, const double* dv, int first,
                           double* r64, CgScalars* outer, double tol, int lanes, hipStream_t s);
void launch_batch_accumulate(const BatchCg& a, const float* xf, double* dv, const CgScalars* outer, hipStream_t s);

}  // namespace femgpu
