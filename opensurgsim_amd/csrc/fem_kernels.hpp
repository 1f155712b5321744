// Host-side launchers of the gfx950 kernels (implemented in fem_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>

namespace femgpu {

// Kernel timing: when a start/stop pair is armed, the next launch made through launch_k carries the
// two events in its dispatch packet (hipExtLaunchKernelGGL), so their interval is the kernel's own
// execution, without the queueing gaps that events recorded around a launch would include.
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
    if (g_launch_timing.start) {
        const LaunchTiming t = g_launch_timing;
        g_launch_timing = LaunchTiming{};
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, t.start, t.stop, 0u, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    }
}

// CG launch geometry: a fixed grid of kCgBlocks blocks of kCgThreads threads (8 waves per SIMD on
// 256 CUs); grid-stride loops cover any size, and every grid reduction has kCgBlocks partials.
constexpr int kCgBlocks = 512;
constexpr int kCgThreads = 1024;
constexpr int kRowGroup = 8;      // lanes per block row in the SpMV
constexpr int kXcds = 8;          // accelerator complex dies per MI355X

struct CgScalars {
    double rz[2];       // r . z, ping-ponged by iteration parity
    double bb;          // |b|^2
    double threshold;   // max(tol^2 |b|^2, realmin)
    double rr;          // |r|^2 at the last test
    double err;         // sqrt(|r|^2 / |b|^2) at exit
    int done;           // solve finished (converged or capped): 0 while running, 1 from every writer but the
                        // one-kernel CG, which writes 2 + k from its launch k (see cg_fused_kernel)
    int converged;
    int iters;          // Eigen's iteration count
    int pad;
};

// Staged record sizes (doubles): an off-diagonal block contribution (9 values + pad, so records start on
// 16-byte boundaries), a diagonal one (lower triangle), a node row of the right-hand side (3 + pad).
constexpr int kPairRec = 10;
constexpr int kDiagRec = 6;
constexpr int kRhsRec = 4;
// Elements per workgroup of the tetrahedron element kernel (256 threads, four lanes per element): the group
// whose contributions are summed in LDS before they are staged.
constexpr int kElemGroup = 64;

struct TetElementArgs {
    int num_elements = 0;
    const int4* conn = nullptr;
    const double* x0 = nullptr;
    const double* xn = nullptr;     // state the element is evaluated at (Newton estimate)
    const double* vn = nullptr;
    const double* xs = nullptr;     // start-of-step state
    const double* vs = nullptr;
    const double4* mat = nullptr;   // lambda, mu, rho, E
    double dt = 0.0;
    int write_matrix = 0;  // 0: only the rhs record
    int general = 0;       // 0: xn == xs and vn == vs (first Newton iteration)
    // lower-block contributions staged by block, one record per element group touching the block, in group
    // order: kPairRec records of the off-diagonal blocks, then kDiagRec records of the diagonal blocks
    double* rec = nullptr;
    double* rhs_rec = nullptr;  // rhs rows staged by node, one kRhsRec record per element group touching the node
    // Per element group g (kElemGroup consecutive elements): records q in [grp_q[g], grp_q[g + 1]) go to offset
    // q_out[q] of rec and sum the LDS contributions src16[q_src[q] .. q_src[q + 1]) (element in group * 10 +
    // slot; slots 0..5 the pairs (0,1), (0,2), (0,3), (1,2), (1,3), (2,3), 6..9 the diagonal blocks), in element
    // order; node rows likewise (grp_n, n_out into rhs_rec, n_src / nsrc16 = element in group * 4 + corner).
    const int* grp_q = nullptr;
    const int* q_out = nullptr;
    const int* q_src = nullptr;
    const uint16_t* src16 = nullptr;
    const int* grp_n = nullptr;
    const int* n_out = nullptr;
    const int* n_src = nullptr;
    const uint16_t* nsrc16 = nullptr;
    int num_groups = 0;
    int* flag_counts = nullptr;  // [3]: det R invalid, G singular, eigen polar path
    const int* elem_inst = nullptr;     // batched instances: instance of each element (null: one dt)
    const double* inst_dt = nullptr;    // time step per instance
    // Operator and right-hand side by integration scheme:
    //   0: A = M (1/dt + alpha_M) + K (dt + alpha_K), rhs row f - K ((dt + alpha_K) vn - (xn - xs))
    //      - alpha_M M vn - M (vn - vs) / dt (implicit Euler with Rayleigh damping D = alpha_M M + alpha_K K);
    //   1: A = M / dt, rhs row f - alpha_K K vn - alpha_M M vn (explicit schemes, Runge-Kutta 4);
    //   2: A = K, rhs row as in 1 (static).
    int mode = 0;
    double ray_mass = 0.0, ray_stiff = 0.0;
    int lumped = 0;                     // lumped mass (row sums on the diagonal blocks)
};

void launch_tet_elements(int kind, const TetElementArgs& a, hipStream_t s);

// The operator: 3x3-block CSR over the symmetric pattern; block k holds its 9 values row-major at
// [9 k, 9 k + 9).
struct Operator {
    int num_nodes;
    int64_t plane;       // number of blocks
    const int* rowptr;   // block CSR
    const int* colidx;
    const int* diagblk;  // block of the diagonal of each row (-1 for a node in no element)
};

// Assembly of the staged contributions into the FP64 operator (and the FP32 copy when S32 != null):
// num_off off-diagonal lower blocks (9 lanes each), then num_diag diagonal ones (6 lanes each).
void launch_assemble(int num_off, int num_diag, const int4* lb_info, const int* ctr_ptr, int nctr_off,
                     const double* stage, double* S64, float* S32, hipStream_t s);
// With damp_d (diag(M) per DOF): the Rayleigh mass damping force -damp_alpha diag(M) v (v = damp_v) added per DOF
// instead of by the elements.
// With node_inst and inst_div (batched instances), each instance's rows are divided by inst_div of the instance.
void launch_rhs(int num_nodes, const int* node_ptr, const double* stage, const double* fg, const double* fext,
                const uint8_t* bcmask, double* b, hipStream_t s, double div = 1.0, const double* damp_d = nullptr,
                const double* damp_v = nullptr, double damp_alpha = 0.0, const int* node_inst = nullptr,
                const double* inst_div = nullptr);
// Localized external stiffness and damping folded into assembled blocks: S[9 blk[e] + k] += add[9 e + k]
// (the FP32 copy rewritten from the FP64 sum when S32 != null).
void launch_ext_operator(int count, const int* blk, const double* add, double* S64, float* S32, hipStream_t s);
// Localized external rows of the right-hand side: b[row_dof[r]] += (f[r] + sum_p kval[p] w[col_dof[p]]) / div
// (div_by_dt: divided by row_dt[r] instead), w = (xn - xs) - vn row_dt[r] when general, -vn row_dt[r] otherwise.
void launch_ext_rhs(int rows, const int* row_dof, const int* row_ptr, const int* col_dof, const double* kval,
                    const double* fval, const double* row_dt, const double* xn, const double* xs, const double* vn,
                    int general, double div, double* b, hipStream_t s, int div_by_dt = 0);
// State update by scheme: 0 (implicit Euler, modified explicit Euler): v += dv, x = xs + dt v;
// 1 (explicit Euler): x = xs + dt v, v += dv; 2 (static): x = xs + dv, v = 0. With node_inst (batched instances):
// dt per instance, inactive instances left as they are, non-finite states flagged per instance.
void launch_update_scheme(int n, int mode, const double* dv, const double* xs, double* x, double* v, double dt,
                          int* invalid, hipStream_t s, const int* node_inst = nullptr, const double* inst_dt = nullptr,
                          const int* inst_active = nullptr, int* inst_invalid = nullptr);
// Runge-Kutta 4 stage state: xo = x + kv dt / div, vo = v + ka dt / div (dt per instance with node_inst).
void launch_rk4_stage(int n, const double* x, const double* v, const double* kv, const double* ka, double dt,
                      double div, double* xo, double* vo, hipStream_t s, const int* node_inst = nullptr,
                      const double* inst_dt = nullptr);
// Runge-Kutta 4 update: x += (v + k4v + 2 (k2v + k3v)) dt / 6, v += (a1 + a4 + 2 (a2 + a3)) dt / 6; batched
// instances as launch_update_scheme.
void launch_rk4_final(int n, double* x, double* v, const double* k2v, const double* k3v, const double* k4v,
                      const double* a1, const double* a2, const double* a3, const double* a4, double dt, int* invalid,
                      hipStream_t s, const int* node_inst = nullptr, const double* inst_dt = nullptr,
                      const int* inst_active = nullptr, int* inst_invalid = nullptr);

// Jacobi-preconditioned CG (Eigen recurrence) on the operator: one fused kernel per iteration, or the
// operator kernel + vector kernel pair.
struct CgArgs {
    Operator op;
    int n;              // 3 * num_nodes
    const void* S;      // TV operator values
    bool fp32;          // TV = float (FP64 reductions)
    void* x;
    void* r;
    void* z;
    void* p[2];         // p of even / odd iterations
    void* q;
    void* dinv;
    double* partial;    // [4 * kCgBlocks]
    CgScalars* sc;
    double tol;
    double floor_tol = 0.0;  // inner solve of the refinement: absolute floor relative to the outer |b| (0: none)
    // fused iteration (one kernel per iteration): second buffers of r, z, q and both of w = dinv q,
    // selected by iteration parity; 7 partial sums per block
    void* r1 = nullptr;
    void* z1 = nullptr;
    void* q1 = nullptr;
    void* w[2] = {nullptr, nullptr};
    double* part7 = nullptr;  // [2][kFusedSums * kCgBlocks], by iteration parity
    void* binv = nullptr;     // FP32 one-kernel iteration: 3x3 block-diagonal preconditioner, 9 floats per node
};
constexpr int kFusedSums = 7;

// One fused CG iteration k (operator, vector update of k - 1, all reductions); see fem_kernels.hip.
void launch_cg_fused(const CgArgs& a, int k, int max_iter, hipStream_t s);
// x = 0, r = (TV) b, z = dinv r and the solve's scalars; skipped when skip && skip->done.
void launch_cg_init(const CgArgs& a, const double* b, const CgScalars* skip, hipStream_t s);
// FP64: the same from the guess a.x (not zeroed): r = b - S x, z = dinv r; the solve's scalars as Eigen's
// solveWithGuess sets them (the tolerance relative to |b|, no iteration when |r| is already below it).
void launch_cg_init_guess(const CgArgs& a, const double* b, hipStream_t s);
// The same for the FP32 one-kernel iteration with the 3x3 block-diagonal preconditioner: a.binv gets
// blockdiag(S)^-1 (9 floats per node), z = binv r.
void launch_cg_init_block(const CgArgs& a, const double* b, const CgScalars* skip, hipStream_t s);
// Iteration k: convergence test of k - 1, p = z + beta p, q = S p, p . q.
void launch_cg_apply(const CgArgs& a, int k, int max_iter, hipStream_t s);
// Iteration k: alpha, x += alpha p, r -= alpha q, z = dinv r, r . r and r . z.
void launch_cg_vector(const CgArgs& a, int k, hipStream_t s);

// Scalars of a solve from the partials of an init kernel (|b|^2 in [0, B), r.z in [B, 2B)).
void launch_cg_init2(const double* partial, CgScalars* sc, const CgScalars* skip, double tol, hipStream_t s,
                     double floor_tol = 0.0);
// Fraction of the outer tolerance an inner FP32 solve of the refinement reduces the outer residual to: its
// threshold is max(inner_tol |r_s|, kInnerFloor tol |b|), so a sweep that starts near the outer tolerance
// stops after the few iterations it needs.
constexpr double kInnerFloor = 0.5;
// *out = ||dv - prev||_inf and prev = dv (partial: kCgBlocks doubles of scratch).
void launch_newton_variation(int64_t n, const double* dv, double* prev, double* partial, double* out,
                             hipStream_t s);
// The vector part of iteration k on raw buffers (used by the element-by-element operators).
void launch_cg_vector_raw(bool fp32, int n, int k, void* x, void* r, void* z, const void* p, const void* q,
                          const void* dinv, double* partial, const CgScalars* sc, hipStream_t s);

// FP64 iterative refinement around the FP32 CG: r64 = b - S64 dv (first: r64 = b) and the outer test.
void launch_residual64(const Operator& op, const double* S64, const double* b, const double* dv, int first,
                       double* r64, double* partial, CgScalars* outer, CgScalars* inner, double tol,
                       hipStream_t s);
// dv += x unless the outer test passed; records the inner iteration count in *iters_out (may be null).
// FP32 refinement on the one-kernel CG path: sweep s's residual, correction and inner-solve start in one
// launch (dv_s = dv_{s-1} + x_{s-1} into dv_out, r64 = b - S64 dv_s, x_out = 0, r = (float) r64, z = B r;
// B formed when first), and the outer test with the inner scalars in one more.
void launch_refine_sweep(const Operator& op, const double* S64, const float* S32, const double* b,
                         const double* dv_in, const float* x_in, double* dv_out, int first, float* x_out, float* r,
                         float* z, float* binv, double* partial, const CgScalars* outer, hipStream_t s);
void launch_refine_check_init(const double* partial, int sweep, double tol, double inner_tol, CgScalars* outer,
                              CgScalars* inner, int* refine_iters, int nrec, hipStream_t s);
void launch_refine_accumulate(int n, const float* xf, double* dv, const CgScalars* outer, const CgScalars* inner,
                              int* iters_out, hipStream_t s);

// State update: v += dv; x = xs + dt v (Math/OdeSolverEulerImplicit.cpp:98-99). Sets *invalid on non-finite.
// v += dv, x = xs + dt v with dt per node's instance when node_inst is given; flags non-finite values.
// Batched: instances with inst_active[i] == 0 are left unchanged; a non-finite value sets inst_invalid[i].
void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s, const int* node_inst, const double* inst_dt, const int* inst_active = nullptr,
                   int* inst_invalid = nullptr);
void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s);

// Full element outputs for tests: R (9), K ((3nn)^2) row-major, f (3nn); FP64.
void launch_eval_elements(int kind, int first, int count, const int4* conn, const double* x0, const double* x,
                          const double4* mat, double* R, double* K, double* f, hipStream_t s);

}  // namespace femgpu
