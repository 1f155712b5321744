// Host-side launchers of the gfx950 kernels (implemented in fem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace femgpu {

// Per-element record written by the element kernels and read by the assembly kernel:
// slots [0, 54): the 6 node-pair blocks (pair order (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)) of
// A_e = M_e / dt + dt K_e, each oriented so that its row node has the larger global id, row-major;
// slots [54, 78): the lower triangles (00,10,11,20,21,22) of the 4 diagonal blocks.
constexpr int kTetRec = 80;   // padded
constexpr int kTetPairSlots = 54;

// Number of blocks of the fixed CG grid (partials per reduction).
constexpr int kCgBlocks = 512;
constexpr int kCgThreads = 256;

struct CgScalars {
    double rz[2];       // r . z, ping-ponged by iteration parity
    double bb;          // |b|^2
    double threshold;   // max(tol^2 |b|^2, realmin)
    double last_rr;     // |r|^2 at the last convergence test
    double err;         // sqrt(|r|^2 / |b|^2) at exit
    int done;           // solve finished (converged or capped)
    int converged;
    int iters;          // Eigen's iteration count
    int pad;
};

struct TetElementArgs {
    int num_elements;
    const int4* conn;
    const double* x0;
    const double* xn;     // state the element is evaluated at (Newton estimate)
    const double* vn;
    const double* xs;     // start-of-step state
    const double* vs;
    const double4* mat;   // lambda, mu, rho, unused
    double dt;
    int write_matrix;     // 0: only the rhs record
    void* rec;            // TK[num_elements][kTetRec]
    double* rhs_rec;      // double[num_elements][12]
    int* flag_counts;     // [3]: det R invalid, G singular, eigen polar path
};

void launch_tet_elements(int kind, bool fp32, const TetElementArgs& a, hipStream_t s);

// Assembles the symmetric (lower-mirrored) BSR operator from the element records.
void launch_assemble(bool fp32, int num_lower, const int* lower_blk, const int* mirror_blk, const int* blk_row,
                     const int* blk_col, const int* ctr_ptr, const int* ctr_list, const void* rec,
                     const uint8_t* bcmask, void* S, hipStream_t s);

void launch_rhs(int num_nodes, const int* node_ptr, const int* node_list, const double* rhs_rec, const double* fg,
                const double* fext, const uint8_t* bcmask, double* b, hipStream_t s);

// CG on the BSR operator.
struct CgArgs {
    int num_nodes;
    int n;              // 3 * num_nodes
    const int* rowptr;  // block CSR
    const int* colidx;
    const int* diagblk;
    const void* S;      // TM[9 * nblocks]
    bool fp32;          // TM = TV = float, FP64 reductions
    void* x;            // TV
    void* r;
    void* p;
    void* q;
    void* dinv;
    const void* b;      // TV right-hand side
    double* partial;    // [4 * kCgBlocks]
    CgScalars* sc;
    double tol;
};
void launch_cg_init(const CgArgs& a, hipStream_t s);
void launch_cg_iteration(const CgArgs& a, int k, int max_iter, hipStream_t s);

// FP64 refinement helpers (FP32 mode).
// r64 = b - S dv (S in float, FP64 arithmetic); partial |r64|^2.
void launch_residual64(const CgArgs& a, const double* b, const double* dv, double* r64, hipStream_t s);
// Sets up an inner solve: rf = (float) r64 (sweep > 0) or (float) b (sweep 0); outer scalars.
void launch_refine_prepare(const CgArgs& a, const double* r64, int first, double tol, CgScalars* outer,
                           hipStream_t s);
// dv += (double) xf
void launch_refine_accumulate(int n, const float* xf, double* dv, const CgScalars* outer, hipStream_t s);

// State update: v += dv; x = xs + dt v (Math/OdeSolverEulerImplicit.cpp:98-99). Sets *invalid on non-finite.
void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s);

// Full element outputs for tests: R (9), K ((3nn)^2) row-major, f (3nn); double.
void launch_eval_elements(int kind, int first, int count, const int4* conn, const double* x0, const double* x,
                          const double4* mat, double* R, double* K, double* f, hipStream_t s);

}  // namespace femgpu
