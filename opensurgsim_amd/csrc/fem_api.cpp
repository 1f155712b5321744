// C ABI of libfem_gpu.so: context, mesh set-up (pattern, assembly maps) and the implicit step.
//
// Host set-up restates (paths relative to the reference root):
//   FemElement::initialize material asserts         Physics/FemElement.cpp:33-43
//   corotational V invertibility assert             Physics/Fem3DElementCorotationalTetrahedron.cpp:57-72
//   lumped node masses                              Physics/FemRepresentation.cpp:102-117
//   sparsity pattern (full 3x3 node-pair blocks)    Physics/FemRepresentation.cpp:121-132
//   boundary conditions                             Math/OdeState.cpp:115-218
//   implicit Euler orchestration                    Math/OdeSolverEulerImplicit.cpp:55-161,
//                                                   Math/OdeSolverLinearEulerImplicit.cpp:43-64
#include "../../include/fem_gpu.h"
#include "fem_kernels.hpp"
#include "fem_adjoint.hpp"
#include "fem_batch.hpp"
#include "fem_hex.hpp"
#include "fem_lu.hpp"
#include "fem_mcg.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

using namespace femgpu;

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidArg : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHECK(expr)                                                                          \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        HIPCHECK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count) HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    ~DevBuf() { release(); }
};

// Events of a timing loop, destroyed when the scope ends (error paths included), with no event pair left
// pending for the next launch.
struct ScopedEvents {
    std::vector<hipEvent_t> ev;
    explicit ScopedEvents(size_t n) {
        ev.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) {
                release();
                throw HipError("hipEventCreate failed");
            }
            ev.push_back(e);
        }
    }
    ~ScopedEvents() { release(); }
    void release() {
        g_launch_timing = LaunchTiming{};
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
    }
};

// HIP-event timers per phase (resolved at host synchronisation points).
struct PhaseTimer {
    bool enabled = false;
    struct Pair {
        hipEvent_t a, b;
        int phase, tag;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Pair> pending;
    hipEvent_t open_ev = nullptr, close_ev = nullptr;
    int open_phase = 0, open_tag = 0;
    double ms[FEM_NUM_PHASES] = {};
    int64_t launches[FEM_NUM_PHASES] = {};
    hipEvent_t get() {
        if (pool.empty()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // Arms the first kernel launched by the phase's launcher with a start/stop event pair.
    void begin(hipStream_t, int phase, int tag = 0) {
        open_ev = get();
        close_ev = get();
        open_phase = phase;
        open_tag = tag;
        g_launch_timing.start = open_ev;
        g_launch_timing.stop = close_ev;
    }
    void end(hipStream_t) {
        if (g_launch_timing.start) {  // nothing was launched: return the pair
            g_launch_timing = LaunchTiming{};
            pool.push_back(open_ev);
            pool.push_back(close_ev);
            return;
        }
        pending.push_back({open_ev, close_ev, open_phase, open_tag});
    }
    // Call after the stream has been synchronised. SpMV pairs with tag > spmv_last were launched after
    // convergence (no work) and are booked separately.
    void resolve(int spmv_last = 1 << 30) {
        for (auto& p : pending) {
            float t = 0.f;
            HIPCHECK(hipEventElapsedTime(&t, p.a, p.b));
            int ph = p.phase;
            if (ph == FEM_PHASE_SPMV && p.tag > spmv_last) ph = FEM_PHASE_SPMV_SKIPPED;
            ms[ph] += t;
            launches[ph]++;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    // As resolve, with the skipped SpMV launches named by a predicate on the tag.
    template <typename Skipped>
    void resolve_by(Skipped skipped) {
        for (auto& p : pending) {
            float t = 0.f;
            HIPCHECK(hipEventElapsedTime(&t, p.a, p.b));
            int ph = p.phase;
            if (ph == FEM_PHASE_SPMV && skipped(p.tag)) ph = FEM_PHASE_SPMV_SKIPPED;
            ms[ph] += t;
            launches[ph]++;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        for (int i = 0; i < FEM_NUM_PHASES; ++i) {
            ms[i] = 0.0;
            launches[i] = 0;
        }
    }
    ~PhaseTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
};

#define TIMED(ctx, phase, tag, stmt)                                   \
    do {                                                               \
        if ((ctx)->timer.enabled) (ctx)->timer.begin((ctx)->stream, phase, tag); \
        stmt;                                                          \
        if ((ctx)->timer.enabled) (ctx)->timer.end((ctx)->stream);     \
    } while (0)

}  // namespace

struct fem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    fem_options opt;
    bool mesh_ready = false;
    int kind = -1;
    int64_t num_nodes = 0, num_elements = 0, ndof = 0;
    int nn = 4;

    // host copies needed after set-up
    std::vector<int32_t> h_rowptr_b, h_colidx_b;   // block CSR
    std::vector<double> h_mass_per_node;
    // OdeEquation::m_initState (Math/OdeEquation.cpp:51-104): F, M, D, K computed at least once (bits 1, 2, 4,
    // 8), and whether the last global M was lumped (computeFMDK with lumping) or consistent (computeM)
    int ode_init = 0;
    bool m_lumped = false;
    // diag(M) per DOF of the tetrahedral mesh: consistent (element diagonals 2 rho V / 20) and lumped (row sums)
    DevBuf<double> mdiag_cons, mdiag_lump;
    std::vector<double> h_nu, h_rho;  // per element, for fem_set_young_modulus
    DevBuf<double> adj_gx, adj_gv;     // adjoint per-element records [E][12]
    std::vector<uint8_t> h_bcmask;
    std::vector<int64_t> h_bc_dofs;
    bool bc_dirty = true;
    bool fg_dirty = true;

    // device data
    DevBuf<double> x0, x, v, xs, vs, fg, fext, b, dv, dvprev, r64;
    DevBuf<double> dv2;         // FP32 refinement: dv of the odd sweeps (dv_s in dv / dv2 by sweep parity)
    DevBuf<int4> conn4;
    DevBuf<double4> mat;
    DevBuf<double> rec;         // element records
    DevBuf<double> rhs_rec;
    DevBuf<int> lower_blk, mirror_blk, blk_row, colidx_b, rowptr_b, ctr_ptr, node_ptr, node_list, diagblk;
    // per element group (the element kernel's workgroup): its combined block records and node rows (see
    // build_tet_maps); rhs_ptr: the rhs records of each node
    DevBuf<int> grp_q, q_out, q_src, grp_n, n_out, n_src, rhs_ptr;
    DevBuf<uint16_t> src16, nsrc16;
    int64_t num_groups = 0;
    int64_t rhs_len = 0;    // doubles of the rhs staging
    DevBuf<int4> lb_info;           // per lower block: block, mirror block, boundary bits (row DOFs 0-2, column 3-5)
    std::vector<int4> h_lb_info;
    std::vector<int32_t> h_lb_row, h_lb_col;
    DevBuf<double> S64;         // operator values (FP64, 9 per block)
    DevBuf<float> S32;          // FP32 copy (FP32 mode)
    DevBuf<uint8_t> bcmask;
    DevBuf<unsigned char> cg_x, cg_r, cg_z, cg_p0, cg_p1, cg_q, cg_dinv, cg_b;  // CG vectors (TV)
    DevBuf<float> cg_binv;  // FP32 one-kernel CG: 3x3 block-diagonal preconditioner, 9 per node
    DevBuf<unsigned char> cg_x2;  // FP32 refinement: inner solution of the odd sweeps
    DevBuf<unsigned char> cg_r1, cg_z1, cg_q1, cg_w0, cg_w1;                    // fused CG: second buffers
    DevBuf<double> part7;      // [2][kFusedSums * kCgBlocks] block sums of the fused CG kernel, by parity
    DevBuf<double> partial;    // [4 * kCgBlocks] block sums of the grid reductions
    DevBuf<CgScalars> scal;    // [0] inner/FP64 solve, [1] outer refinement
    DevBuf<CgScalars> time_sc; // operator replay timing (fem_time_operator)
    DevBuf<int> counters;      // [0..2] element flags, [3] invalid state, [4] Newton max-norm bits (int pair)
    DevBuf<unsigned long long> nrm;
    int num_lower = 0;
    int num_lower_off = 0;  // off-diagonal lower blocks (they precede the diagonal ones)
    int64_t nctr_off = 0;   // contributions to off-diagonal blocks
    int64_t stage_len = 0;  // doubles of the block contribution staging
    int64_t num_blocks = 0;
    bool have_fext = false;

    // localized external terms of the next step (fem_add_external_generalized_force)
    std::vector<int32_t> h_conn;   // element connectivity, num_elements x nn
    // localized constraint compliance (fem_constraint_compliance): H^T columns, C H^T, H's non-zeros
    DevBuf<double> con_b, con_x, con_val, con_out;
    DevBuf<int64_t> con_idx;
    std::vector<double> h_loc_f;   // force, ndof
    std::map<std::pair<int64_t, int64_t>, std::array<double, 18>> h_loc_blk;  // node block -> K (9), D (9)
    bool loc_pending = false;
    DevBuf<int> ext_blk, ext_row_dof, ext_row_ptr, ext_col_dof;
    DevBuf<double> ext_add, ext_kval, ext_fval, ext_row_dt, fstep;
    int ext_nblk = 0, ext_nrows = 0;
    bool fstep_set = false;
    // hexahedra: the localized external operator rows of the last step (kept for compliance after it)
    DevBuf<int> hx_row, hx_ptr, hx_col;
    DevBuf<double> hx_val;
    DevBuf<float> hx_val32;
    int hx_rows = 0;
    bool ext_has_k = false;  // this step's localized terms include a stiffness
    bool k_has_ext = false;  // the last global K formed (implicit or static step) included one
    bool v_static_zero = false;  // v was zeroed by a static step and not set since

    // Runge-Kutta 4 stage vectors
    DevBuf<double> rk_x, rk_k2v, rk_k3v, rk_k4v, rk_a1, rk_a2, rk_a3;

    // multi-column compliance solves (fem_mcg.hip): [ndof][8] interleaved vectors, per-column state
    DevBuf<double> mcg_b, mcg_x, mcg_r, mcg_z, mcg_p, mcg_q, mcg_dinv, mcg_part;
    DevBuf<McgColumn> mcg_col;
    McgColumn* h_mcg = nullptr;  // pinned [8]
    void* scratch = nullptr;     // device address of this device's allocation-free workspace
    // Direct solver (FEM_LINEAR_SOLVER_LU): band LU of the assembled operator, refactored after each assembly
    DeviceBandLU lu;
    int lu_bw = 0;               // bandwidth of the pattern in DOFs: 3 max|I - J| + 2
    bool lu_current = false;     // lu holds the factors of the current S64
    // The dense compliance kept by the linear implicit scheme with the direct solver (formed on its first step,
    // as the reference's first pass forms it; n x n column-major, boundary rows and columns zero), valid while the
    // factors it came from are; gemv_part: the column-chunk partial products that apply it.
    DevBuf<double> dense_c, gemv_part;
    bool dense_c_valid = false;

    // hexahedra (element-by-element path)
    HexData hex;

    // batched independent instances (fem_set_instances): instance tables, per-instance CG scalars,
    // row-chunk tables chosen by the autotuner, captured iteration graphs
    int num_instances = 0;
    std::vector<int64_t> h_node_off;
    std::vector<double> h_inst_dt;      // empty: every instance uses opt.dt
    DevBuf<int> elem_inst, node_inst, b_chunk_inst, b_chunk_row0, b_chunk_row1, b_inst_chunk, b_inst_maxit;
    DevBuf<double> b_inst_dt, b_partial;
    DevBuf<CgScalars> b_sc, b_outer;
    DevBuf<int> b_kA, b_kV;
    DevBuf<int> b_active, b_invalid;   // [M] per-instance activity / non-finite flag of the last update
    // FP32 + refinement mode: instances for which FP32 CG proved inadequate are solved in FP64 from then on
    // (h_inst_fp64, the FP32 pass's skip flags in b_skip32); b_skip64 flags the instances a step's FP64 pass
    // leaves alone
    std::vector<int32_t> h_inst_fp64;
    DevBuf<int> b_skip32, b_skip64;
    DevBuf<int> b_live;                // compacted chunk list of instances still iterating
    int* h_live = nullptr;             // pinned staging for b_live
    std::vector<int32_t> h_inst_active;
    std::vector<int64_t> h_inst_unconv;   // steps whose solve stopped at the cap, per instance
    CgScalars* h_bsc = nullptr;         // pinned [M]
    int b_num_chunks = 0;
    int b_chunk_rows = 0, b_lanes = 0;  // current table shape
    int b_fixed_rows = 0, b_fixed_lanes = 0;  // 0: autotune
    bool b_tuned = false;
    double b_tune_ms = 0.0;             // tuned time per CG iteration (ms)
    double b_maxit_for = -2.0;          // cg_max_iterations the iteration caps were built for
    double b_dt_for = -1.0;             // opt.dt the per-instance dt table was built for
    int b_predicted[17];
    std::vector<int64_t> h_inst_iters;  // last step, per instance (inner iterations summed over sweeps)
    std::vector<double> h_inst_err;
    std::vector<int32_t> h_inst_conv;
    hipGraph_t b_graph[2] = {nullptr, nullptr};
    hipGraphExec_t b_exec[2] = {nullptr, nullptr};  // 8-iteration chunk per precision
    bool b_use_graphs = true;

    // host-pinned readback
    CgScalars* h_sc = nullptr;
    int* h_counters = nullptr;
    DevBuf<double> nvar;         // Newton-Raphson ||dv - dv_prev||_inf of the last iteration
    double* h_nvar = nullptr;    // pinned copy

    // state
    bool have_operator = false;   // a step assembled (or cached) the system matrix
    bool linear_matrix_ready = false;
    double matrix_dt = 0.0;
    int last_precision = -1;
    int predicted_iters = 16;     // FP64 solve / current inner solve
    int predicted_adj = 16;       // FP64 transpose solve of the adjoint step
    bool solving_adjoint = false; // linear_solve serves fem_adjoint_step (its own iteration predictor)
    bool adj_warm = false;        // the FP64 transpose solve starts from dv (the previous adjoint step's solution)
    int cg_kernels = 0;           // 0, 1: fused single kernel, 2: apply + vector kernels (FEM_CG_KERNELS)
    int predicted_sweep[16];       // FP32 inner solves, per refinement sweep
    int predicted_sweeps = 3;      // FP32 refinement sweeps queued without a host round trip
    DevBuf<int> refine_iters;      // inner iteration count per sweep (-1: no correction applied)
    int* h_refine_iters = nullptr; // pinned copy
    int refine_iters_cap = 0;
    int predicted_hex[17];         // hexahedral path: FP64 solve, then per refinement sweep
    fem_stats stats;
    PhaseTimer timer;
    int timer_every = 0;           // instrument one step in timer_every (0: off)
    int64_t timer_count = 0;
};

namespace {

int fail(fem_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <typename F>
int guarded(fem_ctx* c, F&& f) {
    if (!c) return FEM_ERR_INVALID;
    try {
        return f();
    } catch (const InvalidArg& e) {
        return fail(c, FEM_ERR_INVALID, e.what());
    } catch (const HipError& e) {
        return fail(c, FEM_ERR_HIP, e.what());
    } catch (const std::exception& e) {
        return fail(c, FEM_ERR_INVALID, e.what());
    }
}

// Signed tet volume as Physics/Fem3DElementTetrahedron.cpp:218-235.
double det3v(const double* a, const double* b, const double* c) {
    return a[0] * b[1] * c[2] + a[2] * b[0] * c[1] + a[1] * b[2] * c[0] - a[2] * b[1] * c[0] - a[1] * b[0] * c[2] -
           a[0] * b[2] * c[1];
}
double tet_volume(const double* p0, const double* p1, const double* p2, const double* p3) {
    return (det3v(p1, p2, p3) - det3v(p0, p2, p3) + det3v(p0, p1, p3) - det3v(p0, p1, p2)) / 6.0;
}

void check_material(double E, double nu, double rho, int64_t e) {
    auto bad = [&](const char* m) { throw InvalidArg(std::string(m) + " (element " + std::to_string(e) + ")"); };
    if (rho == 0.0) bad("Mass density is not set");
    if (nu == 0.0) bad("Poisson ratio is not set");
    if (E == 0.0) bad("Young modulus is not set");
    if (!(rho > 0.0)) bad("Mass density is invalid, it should be positive");
    if (!(nu > 0.0 && nu < 0.5)) bad("Poisson ratio is invalid, it should be within [0 0.5)");
    if (!(E > 0.0)) bad("Young modulus is invalid, it should be positive");
}

void ensure_bc(fem_ctx* c) {
    if (c->bc_dirty) {
        c->bcmask.upload(c->h_bcmask.data(), c->h_bcmask.size(), c->stream);
        if (!c->h_lb_info.empty()) {
            // boundary bits of every lower block for the assembly
            for (size_t t = 0; t < c->h_lb_info.size(); ++t) {
                const int32_t I = c->h_lb_row[t], J = c->h_lb_col[t];
                int bits = 0;
                for (int k = 0; k < 3; ++k) {
                    if (c->h_bcmask[3 * (size_t)I + k]) bits |= 1 << k;
                    if (c->h_bcmask[3 * (size_t)J + k]) bits |= 8 << k;
                }
                c->h_lb_info[t].z = bits;
            }
            c->lb_info.upload(c->h_lb_info.data(), c->h_lb_info.size(), c->stream);
            HIPCHECK(hipStreamSynchronize(c->stream));  // the host tables may change before the copy ran
        }
        c->bc_dirty = false;
    }
    if (c->fg_dirty) {
        // gravity force m_i g (Physics/FemRepresentation.cpp:641-662)
        std::vector<double> fg(c->ndof, 0.0);
        if (c->opt.gravity_enabled)
            for (int64_t n = 0; n < c->num_nodes; ++n)
                for (int k = 0; k < 3; ++k) fg[3 * n + k] = c->opt.gravity[k] * (1.0 * c->h_mass_per_node[n]);
        c->fg.upload(fg.data(), fg.size(), c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->fg_dirty = false;
    }
}

void reset_batch_graphs(fem_ctx* c) {
    for (int i = 0; i < 2; ++i) {
        if (c->b_exec[i]) (void)hipGraphExecDestroy(c->b_exec[i]);
        if (c->b_graph[i]) (void)hipGraphDestroy(c->b_graph[i]);
        c->b_exec[i] = nullptr;
        c->b_graph[i] = nullptr;
    }
}

void alloc_precision_buffers(fem_ctx* c) {
    int prec = c->opt.precision;
    if (prec == c->last_precision) return;
    const bool fp32 = prec == FEM_PRECISION_FP32_REFINED;
    size_t tv = fp32 ? 4 : 8;
    size_t n = (size_t)c->ndof;
    if (c->kind != FEM_ELEMENT_HEX) {
        if (!c->rec.p) c->rec.alloc((size_t)c->stage_len);
        if (!c->S64.p) c->S64.alloc((size_t)c->num_blocks * 9);
        if (fp32) c->S32.alloc((size_t)c->num_blocks * 9);
        else c->S32.release();
    }
    // batched instances in the FP32 mode solve some instances in an FP64 pass on the same vectors
    const size_t tvb = c->num_instances > 0 ? 8 : tv;
    for (auto* b : {&c->cg_x, &c->cg_r, &c->cg_z, &c->cg_p0, &c->cg_p1, &c->cg_q, &c->cg_dinv, &c->cg_b})
        b->alloc(n * tvb);
    if (c->kind != FEM_ELEMENT_HEX)
        for (auto* b : {&c->cg_r1, &c->cg_z1, &c->cg_q1, &c->cg_w0, &c->cg_w1}) b->alloc(n * tv);
    if (c->kind != FEM_ELEMENT_HEX && fp32) {
        c->cg_binv.alloc(3 * n);
        c->cg_x2.alloc(n * tv);
        if (c->dv2.n != n) c->dv2.alloc(n);
    } else {
        c->cg_binv.release();
        c->cg_x2.release();
    }
    if (!c->part7.p) c->part7.alloc(2 * (size_t)kFusedSums * kCgBlocks);
    c->last_precision = prec;
    c->linear_matrix_ready = false;
    reset_batch_graphs(c);
}

Operator op_of(fem_ctx* c) {
    Operator o;
    o.num_nodes = (int)c->num_nodes;
    o.plane = c->num_blocks;
    o.rowptr = c->rowptr_b.p;
    o.colidx = c->colidx_b.p;
    o.diagblk = c->diagblk.p;
    return o;
}

CgArgs cg_args(fem_ctx* c, bool fp32, double tol) {
    CgArgs a;
    a.op = op_of(c);
    a.n = (int)c->ndof;
    a.S = fp32 ? (const void*)c->S32.p : (const void*)c->S64.p;
    a.fp32 = fp32;
    a.x = c->cg_x.p;
    a.r = c->cg_r.p;
    a.z = c->cg_z.p;
    a.p[0] = c->cg_p0.p;
    a.p[1] = c->cg_p1.p;
    a.q = c->cg_q.p;
    a.dinv = c->cg_dinv.p;
    a.partial = c->partial.p;
    a.sc = c->scal.p;
    a.tol = tol;
    a.r1 = c->cg_r1.p;
    a.z1 = c->cg_z1.p;
    a.q1 = c->cg_q1.p;
    a.w[0] = c->cg_w0.p;
    a.w[1] = c->cg_w1.p;
    a.part7 = c->part7.p;
    a.binv = c->cg_binv.p;
    return a;
}

// One kernel per CG iteration (cg_fused_kernel) unless FEM_CG_KERNELS=split selects the operator kernel +
// vector kernel pair (same recurrence, two launches and one more pass over the vectors per iteration).
bool use_fused_cg(const fem_ctx* c, bool fp32) {
    (void)fp32;
    return c->cg_kernels != 2;
}

// Fused path: the init kernel writes r_0, z_0 where the fused kernel of iteration 0 reads them (odd
// buffers).
void launch_solver_init(fem_ctx* c, const CgArgs& a, const double* b, const CgScalars* skip) {
    if (!use_fused_cg(c, a.fp32)) {
        launch_cg_init(a, b, skip, c->stream);
        return;
    }
    CgArgs t = a;
    t.r = a.r1;
    t.z = a.z1;
    // FP32 inner solves of the refinement: block-diagonal preconditioner (the outer FP64 test sets the
    // accuracy); FP64: Eigen's diagonal preconditioner, from zero or (a warm adjoint solve) from the guess in x
    if (a.fp32) launch_cg_init_block(t, b, skip, c->stream);
    else if (c->adj_warm && c->solving_adjoint) launch_cg_init_guess(t, b, c->stream);
    else launch_cg_init(t, b, skip, c->stream);
}

// Iteration k of an initialised solve (timer tag `tag`): the fused kernel, or the operator + vector pair.
void launch_iteration(fem_ctx* c, const CgArgs& a, int k, int max_iter, int tag) {
    if (use_fused_cg(c, a.fp32)) {
        TIMED(c, FEM_PHASE_SPMV, tag, launch_cg_fused(a, k, max_iter, c->stream));
        return;
    }
    TIMED(c, FEM_PHASE_SPMV, tag, launch_cg_apply(a, k, max_iter, c->stream));
    if (k < max_iter) TIMED(c, FEM_PHASE_CG_VECTOR, tag, launch_cg_vector(a, k, c->stream));
}

int64_t max_iters(fem_ctx* c) {
    int64_t m = c->opt.cg_max_iterations < 0 ? 2 * c->ndof : c->opt.cg_max_iterations;
    return std::min<int64_t>(m, std::numeric_limits<int>::max() - 1);
}

// Runs the CG iterations of an initialised solve: launches a predicted number of iterations, then
// polls the device 'done' flag between chunks. Returns the iteration count.
int run_cg_iterations(fem_ctx* c, const CgArgs& a, int max_iter) {
    int& pred = c->solving_adjoint ? c->predicted_adj : c->predicted_iters;
    int k = 0;
    // iteration k = convergence test of k-1 + p update + operator application, then vector updates;
    // the last launch (k == max_iter) only tests.
    auto iteration = [&](int kk) { launch_iteration(c, a, kk, max_iter, kk); };
    // iteration k tests iteration k - 1: a converged solve with iters = i needs launches 0 .. i + 1
    int first = std::max(1, std::min(pred + 1, max_iter + 1));
    for (; k < first; ++k) iteration(k);
    for (;;) {
        HIPCHECK(hipMemcpyAsync(c->h_sc, a.sc, sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->h_sc->done || k > max_iter) break;
        int chunk = std::min(std::max(2, pred / 8), max_iter + 1 - k);
        for (int j = 0; j < chunk; ++j, ++k) iteration(k);
    }
    int it = c->h_sc->iters;
    if (c->timer.enabled) c->timer.resolve(it);
    pred = it + 1;
    return it;
}

// Linear solve S dv = b (b in c->b, FP64). FP64: Eigen-CG on the FP64 operator. FP32: FP32 CG inside
// FP64 iterative refinement on the implicit-step residual. Result in c->dv.
// The band LU of the current operator S64 (factored once per assembly; the linear schemes keep the first one).
void ensure_lu(fem_ctx* c, bool may_allocate = true) {
    if (c->lu_current) return;
    if (c->kind == FEM_ELEMENT_HEX || c->num_instances > 0 || !c->S64.p)
        throw InvalidArg("the direct solver needs the assembled operator of a tetrahedral mesh of one instance");
    DeviceBandLU& lu = c->lu;
    const size_t need = band_bytes(c->ndof, c->lu_bw);
    if (lu.capacity * sizeof(double) < need) {
        if (!may_allocate) throw InvalidArg("no band LU storage allocated yet: step once with the LU solver first");
        size_t free_b = 0, total_b = 0;
        HIPCHECK(hipMemGetInfo(&free_b, &total_b));
        if (need + (size_t(1) << 30) > free_b + lu.capacity * sizeof(double))
            throw InvalidArg("the banded LU of " + std::to_string(c->ndof) + " DOFs (bandwidth " +
                             std::to_string(c->lu_bw) + ") needs " + std::to_string(need >> 20) + " MiB; " +
                             std::to_string(free_b >> 20) + " MiB free");
        if (lu.ab) HIPCHECK(hipFree(lu.ab));
        lu.ab = nullptr;
        lu.capacity = 0;
        HIPCHECK(hipMalloc(&lu.ab, need));
        lu.capacity = need / sizeof(double);
    }
    lu.n = c->ndof;
    lu.bw = c->lu_bw;
    lu.ld = 2 * (int64_t)c->lu_bw + 1;
    lu.inv = lu.ab + lu.n * lu.ld;
    launch_band_fill(lu, c->num_blocks, c->blk_row.p, c->colidx_b.p, c->S64.p, c->stream);
    launch_band_factor(lu, c->stream);
    c->lu_current = true;
}

// dv = A^-1 b through the band LU (LinearSparseSolveAndInverseLU::solve, Math/LinearSparseSolveAndInverse.cpp:92-97),
// or, when the dense compliance of the same factors is kept, dv = C b: the same solution to rounding (C has the
// boundary rows and columns zeroed; b's boundary rows are zero, and the band solve returns them zero too).
void lu_linear_solve(fem_ctx* c) {
    if (c->dense_c_valid) {
        TIMED(c, FEM_PHASE_CG_INIT, 0,
              launch_dense_gemv(c->ndof, c->dense_c.p, c->b.p, c->gemv_part.p, c->dv.p, c->stream));
    } else {
        ensure_lu(c);
        TIMED(c, FEM_PHASE_CG_INIT, 0, launch_band_solve(c->lu, 1, c->b.p, c->ndof, c->dv.p, c->ndof, c->stream));
    }
    c->stats.last_cg_iterations = 0;
    c->stats.last_cg_error = 0.0;
    c->stats.converged = 1;
    c->stats.last_refinements = 0;
}

void linear_solve(fem_ctx* c) {
    if (c->opt.linear_solver == FEM_LINEAR_SOLVER_LU) return lu_linear_solve(c);
    const bool fp32 = c->opt.precision == FEM_PRECISION_FP32_REFINED;
    const int max_iter = (int)max_iters(c);
    size_t n = (size_t)c->ndof;
    if (!fp32) {
        CgArgs a = cg_args(c, false, c->opt.cg_tolerance);
        a.x = c->dv.p;
        TIMED(c, FEM_PHASE_CG_INIT, 0, launch_solver_init(c, a, c->b.p, nullptr));
        int it = run_cg_iterations(c, a, max_iter);
        c->stats.last_cg_iterations = it;
        c->stats.total_cg_iterations += it;
        c->stats.last_cg_error = c->h_sc->err;
        c->stats.converged = c->h_sc->converged;
        c->stats.last_refinements = 0;
        return;
    }
    // FP32 CG inside FP64 refinement. Sweep s: FP64 residual of dv and the outer test; unless that passed,
    // an FP32 CG on the residual whose iterations are capped a little above the count sweep s took in the
    // last step (launches past convergence cost a dispatch each; a capped inner solve still leaves its x
    // complete and the outer test decides whether another sweep follows), then dv += x. The
    // predicted number of sweeps goes out back to back with no host round trip; the host reads the outcome
    // once and continues sweep by sweep only when the prediction fell short.
    const int max_ref = std::max(0, c->opt.max_refinements);
    if (c->refine_iters_cap < max_ref + 1) {
        c->refine_iters.alloc((size_t)max_ref + 1);
        if (c->h_refine_iters) (void)hipHostFree(c->h_refine_iters);
        c->h_refine_iters = nullptr;
        HIPCHECK(hipHostMalloc((void**)&c->h_refine_iters, ((size_t)max_ref + 1) * sizeof(int), hipHostMallocDefault));
        c->refine_iters_cap = max_ref + 1;
    }
    CgArgs a = cg_args(c, true, c->opt.inner_tolerance);
    a.floor_tol = kInnerFloor * c->opt.cg_tolerance;
    CgScalars* outer = c->scal.p + 1;
    // One-kernel CG path: sweep s is one launch for dv_s = dv_{s-1} + x_{s-1}, the FP64 residual and the inner
    // solve's start, and one for the outer test with the inner scalars; dv_s and x_s alternate between two
    // buffers by parity (dv_0 = 0 in dv). The operator + vector pair path keeps separate launches.
    const bool fused = use_fused_cg(c, true);
    double* dvb[2] = {c->dv.p, c->dv2.p};
    float* xb[2] = {(float*)c->cg_x.p, (float*)c->cg_x2.p};
    if (!fused) {
        HIPCHECK(hipMemsetAsync(c->dv.p, 0, n * sizeof(double), c->stream));
        HIPCHECK(hipMemsetAsync(c->refine_iters.p, 0xff, (size_t)(max_ref + 1) * sizeof(int), c->stream));
    }
    auto check = [&](int sweep) {
        if (fused) {
            TIMED(c, FEM_PHASE_REFINE, 0,
                  launch_refine_sweep(a.op, c->S64.p, c->S32.p, c->b.p, dvb[(sweep + 1) & 1], xb[(sweep + 1) & 1],
                                      dvb[sweep & 1], sweep == 0, xb[sweep & 1], (float*)a.r1, (float*)a.z1,
                                      (float*)a.binv, c->partial.p, outer, c->stream));
            launch_refine_check_init(c->partial.p, sweep, c->opt.cg_tolerance, c->opt.inner_tolerance, outer, a.sc,
                                     c->refine_iters.p, max_ref + 1, c->stream);
            return;
        }
        TIMED(c, FEM_PHASE_REFINE, 0,
              launch_residual64(a.op, c->S64.p, c->b.p, c->dv.p, sweep == 0, c->r64.p, c->partial.p, outer, a.sc,
                                c->opt.cg_tolerance, c->stream));
    };
    auto inner = [&](int sweep) {
        const int ps = std::min(sweep, 15);
        const int pred = c->predicted_sweep[ps];
        // a margin of one launch for the last predicted sweep only: an earlier inner solve stopped one iteration
        // short leaves a slightly larger residual for the next one, while the last one decides the outer test
        const bool last = sweep + 1 >= c->predicted_sweeps;
        const int cap = (int)std::min<int64_t>(max_iter, std::max(4, last ? pred + 1 : pred));
        if (fused) {
            CgArgs ai = a;
            ai.x = xb[sweep & 1];
            for (int k = 0; k <= cap; ++k) launch_iteration(c, ai, k, cap, (sweep << 20) | k);
            return;
        }
        TIMED(c, FEM_PHASE_CG_INIT, 0, launch_solver_init(c, a, c->r64.p, outer));
        for (int k = 0; k <= cap; ++k) launch_iteration(c, a, k, cap, (sweep << 20) | k);
        TIMED(c, FEM_PHASE_REFINE, 0,
              launch_refine_accumulate((int)n, (const float*)c->cg_x.p, c->dv.p, outer, a.sc,
                                       c->refine_iters.p + sweep, c->stream));
    };
    auto read_back = [&](int upto) {
        HIPCHECK(hipMemcpyAsync(c->h_sc + 1, outer, sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipMemcpyAsync(c->h_refine_iters, c->refine_iters.p, (size_t)upto * sizeof(int),
                                hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    };
    int sweep = 0;
    const int ahead = std::min(max_ref, std::max(1, c->predicted_sweeps));
    for (; sweep < ahead; ++sweep) {
        check(sweep);
        inner(sweep);
    }
    check(sweep);  // the outer test after the last queued correction
    read_back(std::max(1, sweep));
    while (!c->h_sc[1].done && sweep < max_ref) {
        inner(sweep);
        ++sweep;
        check(sweep);
        read_back(sweep);
    }
    // sweeps that applied a correction recorded their inner iteration counts
    int total = 0, sweeps = 0;
    for (int i = 0; i < sweep; ++i) {
        const int it = c->h_refine_iters[i];
        if (it < 0) break;
        total += it + 1;
        sweeps++;
        c->predicted_sweep[std::min(i, 15)] = it + 1;
    }
    c->predicted_sweeps = std::max(1, sweeps);
    if (fused && (sweeps & 1)) std::swap(c->dv.p, c->dv2.p);  // dv_sweeps, the corrected solution, is in dv2
    if (c->timer.enabled) {
        std::vector<int> last(sweep + 1, -1);
        for (int i = 0; i < sweeps; ++i) last[i] = c->h_refine_iters[i];
        c->timer.resolve_by([&](int tag) {
            const int sw = tag >> 20, k = tag & ((1 << 20) - 1);
            return sw >= (int)last.size() || k > last[sw];
        });
    }
    const CgScalars& o = c->h_sc[1];
    c->stats.last_cg_iterations = total;
    c->stats.total_cg_iterations += total;
    c->stats.last_cg_error = o.err;
    c->stats.converged = o.converged;
    c->stats.last_refinements = sweeps;
}

// Element-kernel mode of an integration scheme: 0 implicit Euler (M / dt + D + dt K), 1 explicit schemes
// and Runge-Kutta 4 (M / dt), 2 static (K).
int scheme_mode(int scheme) {
    if (scheme == FEM_SCHEME_STATIC || scheme == FEM_SCHEME_LINEAR_STATIC) return 2;
    if (scheme >= FEM_SCHEME_EULER_EXPLICIT) return 1;
    return 0;
}
// The linear variants keep the first step's system (Math/OdeSolverLinear*.cpp).
bool scheme_is_linear(int scheme) {
    return scheme == FEM_SCHEME_LINEAR_EULER_IMPLICIT || scheme == FEM_SCHEME_LINEAR_STATIC ||
           scheme == FEM_SCHEME_LINEAR_EULER_EXPLICIT || scheme == FEM_SCHEME_LINEAR_EULER_EXPLICIT_MODIFIED ||
           scheme == FEM_SCHEME_LINEAR_RUNGE_KUTTA_4;
}

// The Rayleigh mass damping force of computeF (Physics/FemRepresentation.cpp:396-410), which calls
// addRayleighDampingForce with useGlobalMassMatrix = hasF() (:565-593): once F has been computed, the global M is
// used, and with lumping only its diagonal: the row sums after an implicit step's computeFMDK (which lumps,
// :499-508), the consistent diagonal after an explicit or Runge-Kutta step's computeM (which does not lump,
// :412-421). True when a non-implicit scheme's force takes that diagonal form; otherwise the elements add
// -alpha_M M_e v_e (the element loop, or the global M v, which is the same sum).
bool diag_mass_damping(const fem_ctx* c) {
    return scheme_mode(c->opt.scheme) != 0 && c->opt.mass_lumping && c->opt.rayleigh_mass != 0.0 &&
           (c->ode_init & 1) != 0;
}

// OdeEquation::updateFMDK's bookkeeping (Math/OdeEquation.cpp:71-104) after F (and M or K) were formed: an
// implicit step's computeFMDK sets every bit and leaves M lumped as configured; computeM leaves it consistent.
void mark_fmdk(fem_ctx* c) {
    c->ode_init = 15;
    c->m_lumped = c->opt.mass_lumping != 0;
    c->k_has_ext = c->ext_has_k;
}
void mark_f(fem_ctx* c, bool matrix) {
    c->ode_init |= 1;
    if (!matrix) return;
    if (scheme_mode(c->opt.scheme) == 2) {
        c->ode_init |= 8;  // computeK
        c->k_has_ext = c->ext_has_k;
    } else {
        c->ode_init |= 2;  // computeM
        c->m_lumped = false;
    }
}

// The element kernel's per-element inputs (connectivity, materials, instances), the group tables, the record
// buffers and the element flag counters (every launch of the kernel takes them from here).
void set_staging(const fem_ctx* c, TetElementArgs& a) {
    a.conn = c->conn4.p;
    a.mat = c->mat.p;
    if (c->num_instances > 0) {  // each instance's own time step
        a.elem_inst = c->elem_inst.p;
        a.inst_dt = c->b_inst_dt.p;
    }
    a.rec = c->rec.p;
    a.rhs_rec = c->rhs_rec.p;
    a.grp_q = c->grp_q.p;
    a.q_out = c->q_out.p;
    a.q_src = c->q_src.p;
    a.src16 = c->src16.p;
    a.grp_n = c->grp_n.p;
    a.n_out = c->n_out.p;
    a.n_src = c->n_src.p;
    a.nsrc16 = c->nsrc16.p;
    a.num_groups = (int)c->num_groups;
    a.flag_counts = c->counters.p;
}

// Element kernel at state (xn, vn) (null: the current state); xs, vs: the step's start state (Newton terms).
void tet_elements(fem_ctx* c, bool write_matrix, const double* xs, const double* vs, const double* xn = nullptr,
                  const double* vn = nullptr) {
    TetElementArgs a;
    a.num_elements = (int)c->num_elements;
    a.x0 = c->x0.p;
    a.xn = xn ? xn : c->x.p;
    a.vn = vn ? vn : c->v.p;
    a.xs = xs;
    a.vs = vs;
    a.dt = c->opt.dt;
    a.write_matrix = write_matrix ? 1 : 0;
    a.mode = scheme_mode(c->opt.scheme);
    a.ray_mass = diag_mass_damping(c) ? 0.0 : c->opt.rayleigh_mass;  // the diagonal form goes with the rhs
    a.ray_stiff = c->opt.rayleigh_stiffness;
    a.lumped = (a.mode == 0 && c->opt.mass_lumping) ? 1 : 0;
    a.general = (a.mode == 0 && (xs != a.xn || vs != a.vn)) ? 1 : 0;
    set_staging(c, a);
    TIMED(c, FEM_PHASE_ELEMENT, 0, launch_tet_elements(c->kind, a, c->stream));
}

void assemble_tet(fem_ctx* c) {
    if (c->timer.enabled) c->timer.begin(c->stream, FEM_PHASE_ASSEMBLE);
    float* s32 = c->opt.precision == FEM_PRECISION_FP32_REFINED ? c->S32.p : nullptr;
    launch_assemble(c->num_lower_off, c->num_lower - c->num_lower_off, c->lb_info.p, c->ctr_ptr.p,
                    (int)c->nctr_off, c->rec.p, c->S64.p, s32, c->stream);
    if (c->ext_nblk) launch_ext_operator(c->ext_nblk, c->ext_blk.p, c->ext_add.p, c->S64.p, s32, c->stream);
    if (c->timer.enabled) c->timer.end(c->stream);
    c->lu_current = false;
    c->dense_c_valid = false;
}

// rhs = (staged element rows + gravity + external forces) / div with boundary rows zeroed, plus the step's
// localized external rows; (xn, vn, xs, vs) as for tet_elements.
void rhs_tet(fem_ctx* c, double div = 1.0, const double* xn = nullptr, const double* vn = nullptr,
             const double* xs = nullptr, const double* vs = nullptr) {
    const bool dmd = diag_mass_damping(c);
    const double* dd = dmd ? (c->m_lumped ? c->mdiag_lump.p : c->mdiag_cons.p) : nullptr;
    // batched instances: a divisor other than 1 is the step size (the Runge-Kutta rhs F / dt), each instance's own
    const bool inst_div = c->num_instances > 0 && div != 1.0;
    TIMED(c, FEM_PHASE_RHS, 0, launch_rhs((int)c->num_nodes, c->rhs_ptr.p, c->rhs_rec.p, c->fg.p,
                                          c->have_fext ? c->fext.p : nullptr, c->bcmask.p, c->b.p, c->stream, div, dd,
                                          vn ? vn : c->v.p, c->opt.rayleigh_mass, inst_div ? c->node_inst.p : nullptr,
                                          inst_div ? c->b_inst_dt.p : nullptr));
    if (c->ext_nrows) {
        xn = xn ? xn : c->x.p;
        vn = vn ? vn : c->v.p;
        xs = xs ? xs : xn;
        vs = vs ? vs : vn;
        const int general = (xs != xn || vs != vn) ? 1 : 0;
        launch_ext_rhs(c->ext_nrows, c->ext_row_dof.p, c->ext_row_ptr.p, c->ext_col_dof.p, c->ext_kval.p,
                       c->ext_fval.p, c->ext_row_dt.p, xn, xs, vn, general, div, c->b.p, c->stream, inst_div ? 1 : 0);
    }
}

void clear_external(fem_ctx* c);

// Device lists of the localized external terms for the next step (Physics/Fem3DRepresentation.cpp:133-198 and
// their use in computeFMDK / computeF / computeK, Physics/FemRepresentation.cpp:396-410, 460-477, 527-545):
//  - operator entries: D + dt K (implicit Euler) or K (static) under Eigen's lower-triangle view (strictly
//    lower node blocks and their mirrors; the diagonal block's lower part mirrored), boundary rows and
//    columns left as assembled; tetrahedra add them to the assembled blocks, hexahedra apply them as rows
//    beside the element-by-element operator;
//  - right-hand-side rows: the force, plus K (xn - xs - vn dt) with the full external K (implicit Euler);
//    hexahedra take the force from the dense fstep vector and only the K rows here.
void prepare_external(fem_ctx* c) {
    c->ext_nblk = c->ext_nrows = 0;
    c->hx_rows = 0;
    c->fstep_set = false;
    c->ext_has_k = false;
    if (!c->loc_pending) return;
    const size_t n = (size_t)c->ndof;
    const bool hex = c->kind == FEM_ELEMENT_HEX;
    if (c->h_loc_f.size() != n) c->h_loc_f.assign(n, 0.0);
    const int mode = scheme_mode(c->opt.scheme);
    if (!c->h_loc_blk.empty() && scheme_is_linear(c->opt.scheme)) {
        clear_external(c);  // the rejected terms do not linger into later steps
        throw InvalidArg("external stiffness and damping are not supported by the linear (constant-operator) schemes");
    }
    for (const auto& kv : c->h_loc_blk)
        for (int k = 0; k < 9; ++k) c->ext_has_k = c->ext_has_k || kv.second[k] != 0.0;
    if (hex) {
        c->fstep.alloc(n);
        c->fstep.upload(c->h_loc_f.data(), n, c->stream);
        c->fstep_set = true;
    }
    auto node_dt = [&](int64_t node) {
        if (c->num_instances > 0 && !c->h_inst_dt.empty()) {
            const int i = (int)(std::upper_bound(c->h_node_off.begin(), c->h_node_off.end(), node) -
                                c->h_node_off.begin()) - 1;
            return c->h_inst_dt[i];
        }
        return c->opt.dt;
    };
    const uint8_t* bc = c->h_bcmask.data();
    // operator entries (node I, node J, row r, column q of the 3x3 block, value) under the lower-triangle view
    auto for_each_operator_entry = [&](auto&& emit) {
        if (mode == 1) return;  // the explicit schemes' operator is M / dt alone
        for (const auto& kv : c->h_loc_blk) {
            const int64_t I = kv.first.first, J = kv.first.second;
            if (I < J) continue;
            const double dt = node_dt(I);
            double W[9];
            for (int k = 0; k < 9; ++k) W[k] = mode == 0 ? kv.second[9 + k] + dt * kv.second[k] : kv.second[k];
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) {
                    if (I > J) {
                        if (!bc[3 * I + r] && !bc[3 * J + q]) {
                            emit(I, J, r, q, W[3 * r + q]);
                            emit(J, I, q, r, W[3 * r + q]);
                        }
                    } else if (!bc[3 * I + r] && !bc[3 * I + q]) {
                        emit(I, I, r, q, r >= q ? W[3 * r + q] : W[3 * q + r]);
                    }
                }
        }
    };
    std::vector<int> blk;
    std::vector<double> vals;
    std::vector<int> op_row, op_ptr(1, 0), op_col;
    std::vector<double> op_val;
    if (hex) {
        std::map<std::pair<int64_t, int64_t>, double> ent;
        for_each_operator_entry([&](int64_t I, int64_t J, int r, int q, double w) { ent[{3 * I + r, 3 * J + q}] += w; });
        for (const auto& kv : ent) {
            if (op_row.empty() || op_row.back() != (int)kv.first.first) {
                if (!op_row.empty()) op_ptr.push_back((int)op_col.size());
                op_row.push_back((int)kv.first.first);
            }
            op_col.push_back((int)kv.first.second);
            op_val.push_back(kv.second);
        }
        if (!op_row.empty()) op_ptr.push_back((int)op_col.size());
    } else {
        auto block_of = [&](int64_t I, int64_t J) -> int64_t {
            const int32_t* b = c->h_colidx_b.data() + c->h_rowptr_b[I];
            const int32_t* e = c->h_colidx_b.data() + c->h_rowptr_b[I + 1];
            const int32_t* it = std::lower_bound(b, e, (int32_t)J);
            if (it == e || *it != J) throw InvalidArg("external block outside the mesh pattern");
            return (int64_t)(it - c->h_colidx_b.data());
        };
        std::map<int64_t, std::array<double, 9>> add;
        for_each_operator_entry(
            [&](int64_t I, int64_t J, int r, int q, double w) { add[block_of(I, J)][3 * r + q] += w; });
        for (const auto& kv : add) {
            blk.push_back((int)kv.first);
            vals.insert(vals.end(), kv.second.begin(), kv.second.end());
        }
    }
    std::map<int64_t, std::vector<std::pair<int, double>>> rows;
    if (!hex)
        for (size_t d = 0; d < n; ++d)
            if (c->h_loc_f[d] != 0.0) rows[(int64_t)d];
    if (mode == 0)
        for (const auto& kv : c->h_loc_blk) {
            const int64_t I = kv.first.first, J = kv.first.second;
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q)
                    if (kv.second[3 * r + q] != 0.0) rows[3 * I + r].push_back({(int)(3 * J + q), kv.second[3 * r + q]});
        }
    std::vector<int> row_dof, row_ptr(1, 0), col_dof;
    std::vector<double> kval, fval, row_dt;
    for (const auto& kv : rows) {
        if (bc[kv.first]) continue;
        row_dof.push_back((int)kv.first);
        fval.push_back(hex ? 0.0 : c->h_loc_f[kv.first]);
        row_dt.push_back(node_dt(kv.first / 3));
        for (const auto& e : kv.second) {
            col_dof.push_back(e.first);
            kval.push_back(e.second);
        }
        row_ptr.push_back((int)col_dof.size());
    }
    if (!blk.empty()) {
        c->ext_blk.alloc(blk.size());
        c->ext_blk.upload(blk.data(), blk.size(), c->stream);
        c->ext_add.alloc(vals.size());
        c->ext_add.upload(vals.data(), vals.size(), c->stream);
    }
    std::vector<float> op_val32(op_val.begin(), op_val.end());
    if (!op_row.empty()) {
        c->hx_row.alloc(op_row.size());
        c->hx_row.upload(op_row.data(), op_row.size(), c->stream);
        c->hx_ptr.alloc(op_ptr.size());
        c->hx_ptr.upload(op_ptr.data(), op_ptr.size(), c->stream);
        c->hx_col.alloc(op_col.size());
        c->hx_col.upload(op_col.data(), op_col.size(), c->stream);
        c->hx_val.alloc(op_val.size());
        c->hx_val.upload(op_val.data(), op_val.size(), c->stream);
        c->hx_val32.alloc(op_val32.size());
        c->hx_val32.upload(op_val32.data(), op_val32.size(), c->stream);
    }
    if (!row_dof.empty()) {
        c->ext_row_dof.alloc(row_dof.size());
        c->ext_row_dof.upload(row_dof.data(), row_dof.size(), c->stream);
        c->ext_row_ptr.alloc(row_ptr.size());
        c->ext_row_ptr.upload(row_ptr.data(), row_ptr.size(), c->stream);
        c->ext_fval.alloc(fval.size());
        c->ext_fval.upload(fval.data(), fval.size(), c->stream);
        c->ext_row_dt.alloc(row_dt.size());
        c->ext_row_dt.upload(row_dt.data(), row_dt.size(), c->stream);
        c->ext_col_dof.alloc(std::max<size_t>(1, col_dof.size()));
        c->ext_kval.alloc(std::max<size_t>(1, kval.size()));
        if (!col_dof.empty()) {
            c->ext_col_dof.upload(col_dof.data(), col_dof.size(), c->stream);
            c->ext_kval.upload(kval.data(), kval.size(), c->stream);
        }
    }
    HIPCHECK(hipStreamSynchronize(c->stream));  // host tables are freed on return
    c->ext_nblk = (int)blk.size();
    c->ext_nrows = (int)row_dof.size();
    c->hx_rows = (int)op_row.size();
}

// DeformableRepresentation::afterUpdate (Physics/DeformableRepresentation.cpp:236-253): the localized
// external terms last one step.
void clear_external(fem_ctx* c) {
    if (!c->loc_pending) return;
    std::fill(c->h_loc_f.begin(), c->h_loc_f.end(), 0.0);
    c->h_loc_blk.clear();
    c->loc_pending = false;
    c->ext_nblk = c->ext_nrows = 0;
    c->fstep_set = false;
}

void state_update(fem_ctx* c, const double* xs) {
    TIMED(c, FEM_PHASE_UPDATE, 0,
          launch_update((int)c->ndof, c->dv.p, xs, c->x.p, c->v.p, c->opt.dt, c->counters.p + 3, c->stream));
}

// ------------------------------------------------------------------------------------------
// Batched instances
// ------------------------------------------------------------------------------------------
constexpr int kGraphIters = 8;  // CG iterations per captured graph

// Row-chunk tables: chunks of at most `rows` block rows that never straddle an instance.
void build_batch_tables(fem_ctx* c, int rows, int lanes) {
    const int M = c->num_instances;
    std::vector<int> ci, r0, r1, ic(M + 1, 0);
    for (int i = 0; i < M; ++i) {
        ic[i] = (int)ci.size();
        for (int64_t a = c->h_node_off[i]; a < c->h_node_off[i + 1]; a += rows) {
            ci.push_back(i);
            r0.push_back((int)a);
            r1.push_back((int)std::min<int64_t>(a + rows, c->h_node_off[i + 1]));
        }
    }
    ic[M] = (int)ci.size();
    const int K = (int)ci.size();
    hipStream_t s = c->stream;
    c->b_chunk_inst.alloc(K);
    c->b_chunk_inst.upload(ci.data(), K, s);
    c->b_chunk_row0.alloc(K);
    c->b_chunk_row0.upload(r0.data(), K, s);
    c->b_chunk_row1.alloc(K);
    c->b_chunk_row1.upload(r1.data(), K, s);
    c->b_inst_chunk.alloc(M + 1);
    c->b_inst_chunk.upload(ic.data(), M + 1, s);
    c->b_partial.alloc(4 * (size_t)K);
    c->b_live.alloc(K);
    if (c->h_live) HIPCHECK(hipHostFree(c->h_live));
    c->h_live = nullptr;
    HIPCHECK(hipHostMalloc((void**)&c->h_live, K * sizeof(int), hipHostMallocDefault));
    HIPCHECK(hipStreamSynchronize(s));
    c->b_num_chunks = K;
    c->b_chunk_rows = rows;
    c->b_lanes = lanes;
    reset_batch_graphs(c);
}

// Per-instance iteration caps (Eigen default 2 x instance DOF) and time steps, rebuilt when options change.
void refresh_batch_params(fem_ctx* c) {
    const int M = c->num_instances;
    if (c->b_maxit_for != (double)c->opt.cg_max_iterations) {
        std::vector<int> mx(M);
        for (int i = 0; i < M; ++i) {
            const int64_t d = 3 * (c->h_node_off[i + 1] - c->h_node_off[i]);
            const int64_t m = c->opt.cg_max_iterations < 0 ? 2 * d : c->opt.cg_max_iterations;
            mx[i] = (int)std::min<int64_t>(m, std::numeric_limits<int>::max() - 1);
        }
        c->b_inst_maxit.upload(mx.data(), M, c->stream);
        c->b_maxit_for = (double)c->opt.cg_max_iterations;
    }
    if (c->h_inst_dt.empty() && c->b_dt_for != c->opt.dt) {
        std::vector<double> dt(M, c->opt.dt);
        c->b_inst_dt.upload(dt.data(), M, c->stream);
        c->b_dt_for = c->opt.dt;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
}

BatchCg batch_args(fem_ctx* c, bool fp32) {
    BatchCg a;
    a.t.num_instances = c->num_instances;
    a.t.num_chunks = c->b_num_chunks;
    a.t.chunk_inst = c->b_chunk_inst.p;
    a.t.chunk_row0 = c->b_chunk_row0.p;
    a.t.chunk_row1 = c->b_chunk_row1.p;
    a.t.inst_chunk = c->b_inst_chunk.p;
    a.t.inst_maxit = c->b_inst_maxit.p;
    a.t.inst_active = c->b_active.p;
    a.op = op_of(c);
    a.fp32 = fp32;
    a.S = fp32 ? (const void*)c->S32.p : (const void*)c->S64.p;
    a.x = fp32 ? (void*)c->cg_x.p : (void*)c->dv.p;
    a.r = c->cg_r.p;
    a.z = c->cg_z.p;
    a.p[0] = c->cg_p0.p;
    a.p[1] = c->cg_p1.p;
    a.q = c->cg_q.p;
    a.dinv = fp32 ? (void*)c->cg_binv.p : (void*)c->cg_dinv.p;  // FP32: 3x3 block-diagonal preconditioner
    a.partial = c->b_partial.p;
    a.sc = c->b_sc.p;
    a.kA = c->b_kA.p;
    a.kV = c->b_kV.p;
    return a;
}

// Launches kGraphIters CG iterations: a captured graph replayed (the iteration counters live on the
// device), or plain launches on streams that cannot be captured.
void launch_batch_chunk(fem_ctx* c, const BatchCg& a) {
    const int g = a.fp32 ? 1 : 0;
    hipStream_t s = c->stream;
    if (c->b_use_graphs && s != nullptr) {
        if (!c->b_exec[g]) {
            HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < kGraphIters; ++j) launch_batch_cg_iteration(a, c->b_lanes, s);
            HIPCHECK(hipStreamEndCapture(s, &c->b_graph[g]));
            HIPCHECK(hipGraphInstantiate(&c->b_exec[g], c->b_graph[g], nullptr, nullptr, 0));
        }
        HIPCHECK(hipGraphLaunch(c->b_exec[g], s));
        return;
    }
    for (int j = 0; j < kGraphIters; ++j) launch_batch_cg_iteration(a, c->b_lanes, s);
}

// Runs CG iterations until every instance is done; returns the largest iteration count. While most chunks
// belong to instances still iterating, iterations go out as captured 8-iteration graphs over all chunks;
// once the live instances cover less than half of the chunks, iterations are launched over a compacted list
// of their chunks only (the converged instances' workgroups would only exit).
int run_batch_iterations(fem_ctx* c, const BatchCg& a, int& predicted) {
    const int M = c->num_instances;
    const int K = c->b_num_chunks;
    int launched = 0;
    const int first = std::max(1, predicted + 1);
    while (launched < first) {
        launch_batch_chunk(c, a);
        launched += kGraphIters;
    }
    int cap = 0;
    for (int i = 0; i < M; ++i)
        cap = std::max<int64_t>(cap, c->opt.cg_max_iterations < 0 ? 6 * (c->h_node_off[i + 1] - c->h_node_off[i])
                                                                   : c->opt.cg_max_iterations);
    std::vector<int> ic;
    for (;;) {
        HIPCHECK(hipMemcpyAsync(c->h_bsc, a.sc, M * sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        int live_chunks = 0;
        bool all = true;
        if (ic.empty()) {
            ic.resize(M + 1);
            HIPCHECK(hipMemcpy(ic.data(), c->b_inst_chunk.p, (M + 1) * sizeof(int), hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < M; ++i)
            if (!c->h_bsc[i].done) {
                all = false;
                live_chunks += ic[i + 1] - ic[i];
            }
        if (all || launched > cap + 1) break;
        if (2 * live_chunks > K) {
            const int chunk = std::max(1, std::max(2, predicted / 8) / kGraphIters);
            for (int j = 0; j < chunk; ++j, launched += kGraphIters) launch_batch_chunk(c, a);
        } else {
            int n = 0;
            for (int i = 0; i < M; ++i)
                if (!c->h_bsc[i].done)
                    for (int q = ic[i]; q < ic[i + 1]; ++q) c->h_live[n++] = q;
            HIPCHECK(hipMemcpyAsync(c->b_live.p, c->h_live, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
            BatchCg t = a;
            t.t.live = c->b_live.p;
            t.t.grid = n;
            for (int j = 0; j < 2 * kGraphIters; ++j, ++launched) launch_batch_cg_iteration(t, c->b_lanes, c->stream);
        }
    }
    int mx = 0;
    for (int i = 0; i < M; ++i) mx = std::max(mx, c->h_bsc[i].iters);
    predicted = mx + 1;
    return mx;
}

// Autotuner: time CG iterations over all chunks for each (chunk rows, lanes per row) on the current operator and
// rhs. Shapes within 2 % of the fastest are taken as tied (run to run, close shapes swap places), and the tie goes
// to the fewest rows per chunk, then the fewest lanes per row: once a few long solves remain, their iterations go
// out over the compacted list of their chunks, where smaller chunks leave less idle work (on a batch of 64 uneven
// blocks, 100 steps: 24.9 ms per step with 128-row chunks and 4 lanes, 28.5 ms with 256 rows and 8 lanes, whose
// full iterations were 1.8 % faster). Runs once per context (or after fem_set_batch_tuning(0, 0)).
void tune_batch(fem_ctx* c, bool fp32) {
    if (c->b_fixed_rows > 0) {
        if (c->b_chunk_rows != c->b_fixed_rows || c->b_lanes != c->b_fixed_lanes)
            build_batch_tables(c, c->b_fixed_rows, c->b_fixed_lanes);
        c->b_tuned = true;
        return;
    }
    const int rows_c[] = {64, 128, 256, 512, 1024};
    const int lanes_c[] = {4, 8, 16};
    struct Cand {
        int rows, lanes;
        float ms;  // 8 iterations
    };
    std::vector<Cand> cand;
    ScopedEvents events(2);
    hipEvent_t e0 = events.ev[0], e1 = events.ev[1];
    const bool save = c->b_use_graphs;
    c->b_use_graphs = false;
    for (int rows : rows_c)
        for (int lanes : lanes_c) {
            build_batch_tables(c, rows, lanes);
            BatchCg a = batch_args(c, fp32);
            if (fp32) {
                // FP32 operator on the rhs cast to the CG precision through the init kernel
                launch_batch_cg_init(a, c->b.p, nullptr, c->opt.inner_tolerance, lanes, c->stream);
            } else {
                launch_batch_cg_init(a, c->b.p, nullptr, c->opt.cg_tolerance, lanes, c->stream);
            }
            for (int j = 0; j < 2; ++j) launch_batch_cg_iteration(a, lanes, c->stream);
            HIPCHECK(hipEventRecord(e0, c->stream));
            for (int j = 0; j < 8; ++j) launch_batch_cg_iteration(a, lanes, c->stream);
            HIPCHECK(hipEventRecord(e1, c->stream));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            cand.push_back({rows, lanes, ms});
        }
    c->b_use_graphs = save;
    float best = 1e30f;
    for (const Cand& k : cand) best = std::min(best, k.ms);
    const Cand* pick = nullptr;  // candidates come in increasing rows, then lanes: the first tied one
    for (const Cand& k : cand)
        if (!pick && k.ms <= 1.02f * best) pick = &k;
    build_batch_tables(c, pick->rows, pick->lanes);
    c->b_tune_ms = pick->ms / 8.0;
    c->b_tuned = true;
}

// Per-instance linear solves S_i dv_i = b_i (FP64 CG, or FP32 CG inside FP64 refinement).
void batch_linear_solve(fem_ctx* c) {
    const bool fp32 = c->opt.precision == FEM_PRECISION_FP32_REFINED;
    const int M = c->num_instances;
    refresh_batch_params(c);
    if (!c->b_tuned) tune_batch(c, fp32);
    std::fill(c->h_inst_iters.begin(), c->h_inst_iters.end(), 0);
    if (!fp32) {
        BatchCg a = batch_args(c, false);
        launch_batch_cg_init(a, c->b.p, nullptr, c->opt.cg_tolerance, c->b_lanes, c->stream);
        const int it = run_batch_iterations(c, a, c->b_predicted[0]);
        int64_t total = 0;
        bool conv = true;
        double err = 0.0;
        for (int i = 0; i < M; ++i) {
            c->h_inst_iters[i] = c->h_bsc[i].iters;
            c->h_inst_err[i] = c->h_bsc[i].err;
            c->h_inst_conv[i] = c->h_bsc[i].converged;
            total += c->h_bsc[i].iters;
            conv = conv && c->h_bsc[i].converged;
            err = std::max(err, c->h_bsc[i].err);
        }
        c->stats.last_cg_iterations = it;
        c->stats.total_cg_iterations += total;
        c->stats.last_cg_error = err;
        c->stats.converged = conv;
        c->stats.last_refinements = 0;
        return;
    }
    const size_t n = (size_t)c->ndof;
    HIPCHECK(hipMemsetAsync(c->dv.p, 0, n * sizeof(double), c->stream));
    BatchCg a = batch_args(c, true);
    a.t.inst_skip = c->b_skip32.p;  // instances solved in FP64 (below) are left alone here
    int sweeps = 0, longest = 0;
    std::vector<int> sweep_max(M, 0);  // each instance's longest inner solve of this step
    for (int sweep = 0; sweep <= c->opt.max_refinements; ++sweep) {
        launch_batch_residual(a, c->S64.p, c->b.p, c->dv.p, sweep == 0, c->r64.p, c->b_outer.p, c->opt.cg_tolerance,
                              c->b_lanes, c->stream);
        if (sweep == c->opt.max_refinements) break;
        launch_batch_cg_init(a, c->r64.p, c->b_outer.p, c->opt.inner_tolerance, c->b_lanes, c->stream);
        int& pred = c->b_predicted[1 + std::min(sweep, 15)];
        longest += run_batch_iterations(c, a, pred) + 1;
        for (int i = 0; i < M; ++i)
            if (!c->h_bsc[i].converged || c->h_bsc[i].iters > 0) {
                c->h_inst_iters[i] += c->h_bsc[i].iters;
                sweep_max[i] = std::max(sweep_max[i], c->h_bsc[i].iters);
            }
        HIPCHECK(hipMemcpyAsync(c->h_bsc, c->b_outer.p, M * sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        bool all = true;
        for (int i = 0; i < M && all; ++i) all = c->h_bsc[i].done != 0;
        if (all) break;
        launch_batch_accumulate(a, (const float*)c->cg_x.p, c->dv.p, c->b_outer.p, c->stream);
        sweeps++;
    }
    HIPCHECK(hipMemcpyAsync(c->h_bsc, c->b_outer.p, M * sizeof(CgScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    // FP32 CG is not enough for some operators (ill-conditioned or indefinite ones, typically a collapsing mesh):
    // its inner solves then take many times the iterations of FP64 CG, or reach the cap. An instance whose
    // FP32 refinement stopped short of the outer tolerance, or one of whose inner solves took more than a
    // quarter of its iteration cap (at least 64), is solved in FP64 from the next step on, as the FP64 mode
    // would solve it; one that stopped short is solved in FP64 in this step too. The switch is per instance
    // and lasts until the instances are declared again.
    std::vector<int32_t> solve64(M, 0);
    bool changed = false, any64 = false;
    for (int i = 0; i < M; ++i) {
        if (!c->h_inst_active[i]) continue;
        if (c->h_inst_fp64[i]) {
            solve64[i] = 1;
        } else {
            const int64_t cap = c->opt.cg_max_iterations < 0 ? 6 * (c->h_node_off[i + 1] - c->h_node_off[i])
                                                             : c->opt.cg_max_iterations;
            const bool slow = sweep_max[i] > std::max<int64_t>(64, cap / 4);
            if (!c->h_bsc[i].converged || slow) {
                c->h_inst_fp64[i] = 1;
                changed = true;
            }
            solve64[i] = c->h_bsc[i].converged ? 0 : 1;
        }
        any64 = any64 || solve64[i];
    }
    for (int i = 0; i < M; ++i)
        if (!solve64[i]) {
            c->h_inst_err[i] = c->h_bsc[i].err;
            c->h_inst_conv[i] = c->h_bsc[i].converged;
        }
    if (changed) {
        c->b_skip32.upload(c->h_inst_fp64.data(), M, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
    if (any64) {
        std::vector<int32_t> comp(M);
        for (int i = 0; i < M; ++i) comp[i] = solve64[i] ? 0 : 1;
        c->b_skip64.upload(comp.data(), M, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));  // comp is freed on return
        BatchCg a64 = batch_args(c, false);  // FP64 operator, x = dv, the Jacobi preconditioner
        a64.t.inst_skip = c->b_skip64.p;
        launch_batch_cg_init(a64, c->b.p, nullptr, c->opt.cg_tolerance, c->b_lanes, c->stream);
        longest += run_batch_iterations(c, a64, c->b_predicted[0]);
        for (int i = 0; i < M; ++i)
            if (solve64[i]) {
                c->h_inst_iters[i] += c->h_bsc[i].iters;
                c->h_inst_err[i] = c->h_bsc[i].err;
                c->h_inst_conv[i] = c->h_bsc[i].converged;
            }
    }
    int64_t total = 0;
    bool conv = true;
    double err = 0.0;
    for (int i = 0; i < M; ++i) {
        total += c->h_inst_iters[i];
        conv = conv && c->h_inst_conv[i];
        err = std::max(err, c->h_inst_err[i]);
    }
    c->stats.last_cg_iterations = longest;
    c->stats.total_cg_iterations += total;
    c->stats.last_cg_error = err;
    c->stats.converged = conv;
    c->stats.last_refinements = sweeps;
}

// The batched linear solve of a step, with each instance's count of unconverged solves.
void batch_solve(fem_ctx* c) {
    batch_linear_solve(c);
    for (int i = 0; i < c->num_instances; ++i)
        if (c->h_inst_active[i] && !c->h_inst_conv[i]) c->h_inst_unconv[i]++;
}

// After a batched update: instances whose new state is not finite are deactivated
// (Physics/FemRepresentation.cpp:285-293), the others go on; the batch is inactive only when every instance is.
void batch_deactivate(fem_ctx* c) {
    std::vector<int> inv(c->num_instances);
    HIPCHECK(hipMemcpyAsync(inv.data(), c->b_invalid.p, inv.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    bool changed = false, any = false;
    for (int i = 0; i < c->num_instances; ++i) {
        if (inv[i] && c->h_inst_active[i]) {
            c->h_inst_active[i] = 0;
            changed = true;
        }
        any = any || c->h_inst_active[i];
    }
    if (changed) c->b_active.upload(c->h_inst_active.data(), c->num_instances, c->stream);
    if (!any) HIPCHECK(hipMemsetAsync(c->counters.p + 3, 0xff, sizeof(int), c->stream));
}

// One implicit Euler step of every instance (one Newton iteration); the linear scheme keeps the first step's
// operator (Math/OdeSolverLinearEulerImplicit.cpp:43-64).
int batch_step(fem_ctx* c) {
    const bool linear = c->opt.scheme == FEM_SCHEME_LINEAR_EULER_IMPLICIT;
    if ((c->opt.scheme != FEM_SCHEME_EULER_IMPLICIT && !linear) || c->opt.newton_max_iterations != 1)
        throw InvalidArg("batched instances step with implicit Euler, one Newton iteration");
    const bool reuse = linear && c->linear_matrix_ready && c->matrix_dt == c->opt.dt;
    refresh_batch_params(c);
    TetElementArgs a;
    a.num_elements = (int)c->num_elements;
    a.x0 = c->x0.p;
    a.xn = c->x.p;
    a.vn = c->v.p;
    a.xs = c->x.p;
    a.vs = c->v.p;
    a.dt = c->opt.dt;
    a.write_matrix = reuse ? 0 : 1;
    a.general = 0;
    set_staging(c, a);
    a.ray_mass = c->opt.rayleigh_mass;
    a.ray_stiff = c->opt.rayleigh_stiffness;
    a.lumped = c->opt.mass_lumping ? 1 : 0;
    TIMED(c, FEM_PHASE_ELEMENT, 0, launch_tet_elements(c->kind, a, c->stream));
    if (!reuse) assemble_tet(c);
    rhs_tet(c);
    if (reuse) {
        mark_f(c, false);
    } else {
        mark_fmdk(c);
    }
    batch_solve(c);
    TIMED(c, FEM_PHASE_UPDATE, 0,
          launch_update((int)c->ndof, c->dv.p, c->x.p, c->x.p, c->v.p, c->opt.dt, c->counters.p + 3, c->stream,
                        c->node_inst.p, c->b_inst_dt.p, c->b_active.p, c->b_invalid.p));
    batch_deactivate(c);
    if (linear && !reuse) {
        c->linear_matrix_ready = true;
        c->matrix_dt = c->opt.dt;
    }
    c->stats.steps++;
    return FEM_OK;
}

HexStepArgs hex_args(fem_ctx* c) {
    HexStepArgs a;
    a.opt = &c->opt;
    a.x0 = c->x0.p;
    a.x = c->x.p;
    a.v = c->v.p;
    a.fg = c->fg.p;
    a.fext = c->have_fext ? c->fext.p : nullptr;
    a.fstep = c->fstep_set ? c->fstep.p : nullptr;
    a.ext_rows = c->hx_rows;
    if (c->hx_rows) {
        a.ext_row = c->hx_row.p;
        a.ext_ptr = c->hx_ptr.p;
        a.ext_col = c->hx_col.p;
        a.ext_val = c->hx_val.p;
        a.ext_val32 = c->hx_val32.p;
    }
    a.rhs_rows = c->ext_nrows;
    if (c->ext_nrows) {
        a.rhs_row = c->ext_row_dof.p;
        a.rhs_ptr = c->ext_row_ptr.p;
        a.rhs_col = c->ext_col_dof.p;
        a.rhs_k = c->ext_kval.p;
        a.rhs_f = c->ext_fval.p;
        a.rhs_dt = c->ext_row_dt.p;
    }
    a.bcmask = c->bcmask.p;
    a.mat = c->mat.p;
    a.b = c->b.p;
    a.dv = c->dv.p;
    a.r64 = c->r64.p;
    a.vec[0] = c->cg_x.p;
    a.vec[1] = c->cg_r.p;
    a.vec[2] = c->cg_z.p;
    a.vec[3] = c->cg_p0.p;
    a.vec[4] = c->cg_q.p;
    a.vec[5] = c->cg_dinv.p;
    a.vec[6] = c->dvprev.p;
    a.partial = c->partial.p;
    a.sc = c->scal.p;
    a.h_sc = c->h_sc;
    a.invalid = c->counters.p + 3;
    a.max_iter = (int)max_iters(c);
    a.predicted = c->predicted_hex;
    return a;
}

// Static, explicit Euler, modified explicit Euler and Runge-Kutta 4 steps and their linear variants on the
// assembled tetrahedral operator (Math/OdeSolverStatic.cpp:32-70, Math/OdeSolverEulerExplicit.cpp:32-70,
// Math/OdeSolverEulerExplicitModified.cpp:32-71, Math/OdeSolverRungeKutta4.cpp:32-112 and
// Math/OdeSolverLinear*.cpp). The system matrix is M / dt (K for static); the linear variants keep the first
// step's matrix and solve later steps against it (the compliance of DeformableRepresentation::applyCompliance,
// Physics/DeformableRepresentation.cpp:175-190: boundary rows of the right-hand side and of the solution zero).
// The other schemes on a hexahedral mesh: the same updates as step_other_scheme with the element-by-element
// operator (static: K; explicit and Runge-Kutta 4: M / dt, never lumped, Physics/FemRepresentation.cpp:412-421)
// and the force f(x, v) with the Rayleigh term as the right-hand side. A hexahedron's K and M are constant, so
// the linear variants (first step's operator kept) take the same arithmetic.
void step_other_scheme_hex(fem_ctx* c) {
    const int sc = c->opt.scheme;
    const double dt = c->opt.dt;
    const int n = (int)c->ndof;
    HexStepArgs a = hex_args(c);
    const bool stat = sc == FEM_SCHEME_STATIC || sc == FEM_SCHEME_LINEAR_STATIC;
    hex_set_operator(c->hex, a, stat ? 0.0 : 1.0 / dt, stat ? 1.0 : 0.0, 0, c->stream);
    const double kv = c->opt.rayleigh_stiffness, am = c->opt.rayleigh_mass;
    double* x = c->x.p;
    double* v = c->v.p;
    // the step forms its matrix (computeM or computeK) unless a linear variant keeps the first one
    const bool forms = !(scheme_is_linear(sc) && c->linear_matrix_ready && c->matrix_dt == dt);
    auto damp = [&]() -> const double* {  // computeF's diagonal Rayleigh mass form, when it applies
        return diag_mass_damping(c) ? hex_mass_diagonal(c->hex, c->x0.p, c->mat.p, c->m_lumped, c->stream) : nullptr;
    };
    if (scheme_is_linear(sc) && forms) {
        c->linear_matrix_ready = true;
        c->matrix_dt = dt;
    }
    if (sc == FEM_SCHEME_RUNGE_KUTTA_4 || sc == FEM_SCHEME_LINEAR_RUNGE_KUTTA_4) {
        for (DevBuf<double>* b : {&c->rk_x, &c->rk_k2v, &c->rk_k3v, &c->rk_k4v, &c->rk_a1, &c->rk_a2, &c->rk_a3})
            if (b->n != (size_t)n) b->alloc(n);
        const size_t bytes = (size_t)n * sizeof(double);
        // k1 = (v, M^-1 F(x, v)): f / dt against M / dt
        hex_rhs(c->hex, a, x, v, kv, am, 0, dt, c->stream, damp());
        mark_f(c, forms);
        hex_solve(c->hex, a, c->stats, c->stream);
        HIPCHECK(hipMemcpyAsync(c->rk_a1.p, c->dv.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        auto stage = [&](const double* kvs, const double* kas, double div, double* vstage, double* aout) {
            launch_rk4_stage(n, x, v, kvs, kas, dt, div, c->rk_x.p, vstage, c->stream);
            hex_rhs(c->hex, a, c->rk_x.p, vstage, kv, am, 0, dt, c->stream, damp());
            mark_f(c, false);
            hex_solve(c->hex, a, c->stats, c->stream);
            if (aout) HIPCHECK(hipMemcpyAsync(aout, c->dv.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        };
        stage(v, c->rk_a1.p, 2.0, c->rk_k2v.p, c->rk_a2.p);
        stage(c->rk_k2v.p, c->rk_a2.p, 2.0, c->rk_k3v.p, c->rk_a3.p);
        stage(c->rk_k3v.p, c->rk_a3.p, 1.0, c->rk_k4v.p, nullptr);  // a4 stays in dv
        TIMED(c, FEM_PHASE_UPDATE, 0,
              launch_rk4_final(n, x, v, c->rk_k2v.p, c->rk_k3v.p, c->rk_k4v.p, c->rk_a1.p, c->rk_a2.p, c->rk_a3.p,
                               c->dv.p, dt, c->counters.p + 3, c->stream));
        return;
    }
    hex_rhs(c->hex, a, x, v, kv, am, 0, 1.0, c->stream, damp());
    mark_f(c, forms);
    hex_solve(c->hex, a, c->stats, c->stream);
    const int umode = stat ? 2 : (sc == FEM_SCHEME_EULER_EXPLICIT || sc == FEM_SCHEME_LINEAR_EULER_EXPLICIT) ? 1 : 0;
    TIMED(c, FEM_PHASE_UPDATE, 0, launch_update_scheme(n, umode, c->dv.p, x, x, v, dt, c->counters.p + 3, c->stream));
}

void step_other_scheme(fem_ctx* c) {
    if (c->kind == FEM_ELEMENT_HEX) return step_other_scheme_hex(c);
    const int sc = c->opt.scheme;
    const double dt = c->opt.dt;
    const int n = (int)c->ndof;
    const bool linear = scheme_is_linear(sc);
    const bool reuse = linear && c->linear_matrix_ready && c->matrix_dt == dt;
    double* x = c->x.p;
    double* v = c->v.p;
    // batched instances: every instance steps with its own dt; inactive ones keep their state
    const bool batched = c->num_instances > 0;
    if (batched) refresh_batch_params(c);
    const int* ni = batched ? c->node_inst.p : nullptr;
    const double* idt = batched ? c->b_inst_dt.p : nullptr;
    const int* act = batched ? c->b_active.p : nullptr;
    int* inv = batched ? c->b_invalid.p : nullptr;
    auto solve = [&]() {
        if (batched) batch_solve(c);
        else linear_solve(c);
    };
    if (sc == FEM_SCHEME_RUNGE_KUTTA_4 || sc == FEM_SCHEME_LINEAR_RUNGE_KUTTA_4) {
        for (DevBuf<double>* b : {&c->rk_x, &c->rk_k2v, &c->rk_k3v, &c->rk_k4v, &c->rk_a1, &c->rk_a2, &c->rk_a3})
            if (b->n != (size_t)n) b->alloc(n);
        const size_t bytes = (size_t)n * sizeof(double);
        // k1 = (v, M^-1 F(x, v)): the rhs is F / dt against M / dt
        tet_elements(c, !reuse, x, v);
        if (!reuse) assemble_tet(c);
        rhs_tet(c, dt);
        mark_f(c, !reuse);
        solve();
        HIPCHECK(hipMemcpyAsync(c->rk_a1.p, c->dv.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        auto stage = [&](const double* kv, const double* ka, double div, double* vstage, double* aout) {
            launch_rk4_stage(n, x, v, kv, ka, dt, div, c->rk_x.p, vstage, c->stream, ni, idt);
            tet_elements(c, false, c->rk_x.p, vstage, c->rk_x.p, vstage);
            rhs_tet(c, dt, c->rk_x.p, vstage);
            mark_f(c, false);
            solve();
            if (aout) HIPCHECK(hipMemcpyAsync(aout, c->dv.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        };
        stage(v, c->rk_a1.p, 2.0, c->rk_k2v.p, c->rk_a2.p);
        stage(c->rk_k2v.p, c->rk_a2.p, 2.0, c->rk_k3v.p, c->rk_a3.p);
        stage(c->rk_k3v.p, c->rk_a3.p, 1.0, c->rk_k4v.p, nullptr);  // a4 stays in dv
        TIMED(c, FEM_PHASE_UPDATE, 0,
              launch_rk4_final(n, x, v, c->rk_k2v.p, c->rk_k3v.p, c->rk_k4v.p, c->rk_a1.p, c->rk_a2.p, c->rk_a3.p,
                               c->dv.p, dt, c->counters.p + 3, c->stream, ni, idt, act, inv));
    } else {
        tet_elements(c, !reuse, x, v);
        if (!reuse) assemble_tet(c);
        rhs_tet(c);
        mark_f(c, !reuse);
        solve();
        const int umode = (sc == FEM_SCHEME_STATIC || sc == FEM_SCHEME_LINEAR_STATIC) ? 2
                          : (sc == FEM_SCHEME_EULER_EXPLICIT || sc == FEM_SCHEME_LINEAR_EULER_EXPLICIT) ? 1 : 0;
        TIMED(c, FEM_PHASE_UPDATE, 0,
              launch_update_scheme(n, umode, c->dv.p, x, x, v, dt, c->counters.p + 3, c->stream, ni, idt, act, inv));
    }
    if (linear && !reuse) {
        c->linear_matrix_ready = true;
        c->matrix_dt = dt;
    }
    if (batched) batch_deactivate(c);
}

void keep_dense_compliance(fem_ctx* c);

// One step of the configured scheme (OdeSolver::solve with computeCompliance = false). Returns FEM_OK or
// FEM_INACTIVE.
int step_once(fem_ctx* c) {
    c->stats.last_newton_iterations = 1;  // the Newton-Raphson loop below sets its own count
    if (c->num_instances > 0 && c->opt.scheme < FEM_SCHEME_STATIC) return batch_step(c);
    const double dt = c->opt.dt;
    const size_t n = (size_t)c->ndof;
    const bool linear_scheme = c->opt.scheme == FEM_SCHEME_LINEAR_EULER_IMPLICIT;
    const int newton = std::max(1, c->opt.newton_max_iterations);
    if (c->opt.scheme >= FEM_SCHEME_STATIC) {
        step_other_scheme(c);
    } else if (c->kind == FEM_ELEMENT_HEX) {
        HexStepArgs a = hex_args(c);
        hex_step(c->hex, a, c->stats, c->stream);
        mark_fmdk(c);
    } else if (linear_scheme && c->linear_matrix_ready && c->matrix_dt == dt) {
        // Steps >= 2 of OdeSolverLinearEulerImplicit: element forces only, constant operator.
        tet_elements(c, false, c->x.p, c->v.p);
        rhs_tet(c);
        mark_f(c, false);
        linear_solve(c);
        state_update(c, c->x.p);
    } else if (newton == 1 || linear_scheme) {
        tet_elements(c, true, c->x.p, c->v.p);
        assemble_tet(c);
        rhs_tet(c);
        mark_fmdk(c);
        // the kept compliance is formed before this step's solve, which then applies it like the later steps
        if (linear_scheme && c->opt.linear_solver == FEM_LINEAR_SOLVER_LU) keep_dense_compliance(c);
        linear_solve(c);
        state_update(c, c->x.p);
        if (linear_scheme) {
            c->linear_matrix_ready = true;
            c->matrix_dt = dt;
        }
    } else {
        // Newton-Raphson (Math/OdeSolverEulerImplicit.cpp:84-115): newState starts as the current state.
        HIPCHECK(hipMemcpyAsync(c->xs.p, c->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(c->vs.p, c->v.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipMemsetAsync(c->dvprev.p, 0, n * sizeof(double), c->stream));
        c->stats.last_newton_iterations = 0;
        for (int it = 0; it < newton; ++it) {
            tet_elements(c, true, c->xs.p, c->vs.p);
            assemble_tet(c);
            rhs_tet(c, 1.0, c->x.p, c->v.p, c->xs.p, c->vs.p);
            mark_fmdk(c);
            linear_solve(c);
            state_update(c, c->xs.p);
            // ||dv - dv_prev||_inf < epsilon (default 1e-5, Math/OdeSolverEulerImplicit.cpp:26, 101-112), reduced on
            // the device; only the scalar comes back
            launch_newton_variation((int64_t)n, c->dv.p, c->dvprev.p, c->partial.p, c->nvar.p, c->stream);
            HIPCHECK(hipMemcpyAsync(c->h_nvar, c->nvar.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
            c->stats.last_newton_iterations = it + 1;
            if (*c->h_nvar < c->opt.newton_epsilon) break;
        }
    }
    c->stats.steps++;
    c->have_operator = true;
    return FEM_OK;
}

void build_tet_maps(fem_ctx* c, const int32_t* conn) {
    const int64_t N = c->num_nodes, E = c->num_elements;
    // Node adjacency (every node pair sharing an element, both orders, plus the diagonal).
    std::vector<int64_t> cnt(N + 1, 0);
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) cnt[conn[4 * e + a] + 1] += 4;
    for (int64_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> adj(cnt[N]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) adj[pos[conn[4 * e + a]]++] = conn[4 * e + b];
    c->h_rowptr_b.assign(N + 1, 0);
    c->h_colidx_b.clear();
    c->h_colidx_b.reserve(cnt[N] / 3);
    for (int64_t i = 0; i < N; ++i) {
        auto b = adj.begin() + cnt[i], en = adj.begin() + cnt[i + 1];
        std::sort(b, en);
        auto last = std::unique(b, en);
        c->h_colidx_b.insert(c->h_colidx_b.end(), b, last);
        if (c->h_colidx_b.size() > (size_t)std::numeric_limits<int32_t>::max() / 9)
            throw InvalidArg("pattern too large for int32 scalar indices");
        c->h_rowptr_b[i + 1] = (int32_t)c->h_colidx_b.size();
    }
    adj.clear();
    adj.shrink_to_fit();
    const auto& rp = c->h_rowptr_b;
    const auto& ci = c->h_colidx_b;
    c->num_blocks = (int64_t)ci.size();
    {
        int64_t maxd = 0;
        for (int64_t I = 0; I < N; ++I)
            for (int32_t k = rp[I]; k < rp[I + 1]; ++k) maxd = std::max<int64_t>(maxd, std::llabs((int64_t)ci[k] - I));
        c->lu_bw = (int)(3 * maxd + 2);
    }
    auto find_blk = [&](int32_t I, int32_t J) -> int32_t {
        auto b = ci.begin() + rp[I], en = ci.begin() + rp[I + 1];
        auto it = std::lower_bound(b, en, J);
        return (int32_t)(it - ci.begin());
    };
    std::vector<int32_t> blk_row(c->num_blocks), diag(N, -1);
    std::vector<int32_t> lower_of_blk(c->num_blocks, -1);
    std::vector<int32_t> lower_blk, mirror;
    // Lower blocks: the off-diagonal ones in row-major order, then the diagonal ones in row order (waves of
    // the assembly then see one kind; a diagonal block has about four times the contributions).
    for (int32_t I = 0; I < (int32_t)N; ++I)
        for (int32_t k = rp[I]; k < rp[I + 1]; ++k) {
            blk_row[k] = I;
            int32_t J = ci[k];
            if (J == I) diag[I] = k;
            if (J < I) {
                lower_of_blk[k] = (int32_t)lower_blk.size();
                lower_blk.push_back(k);
                mirror.push_back(find_blk(J, I));
            }
        }
    c->num_lower_off = (int)lower_blk.size();
    for (int32_t I = 0; I < (int32_t)N; ++I)
        if (diag[I] >= 0) {
            lower_of_blk[diag[I]] = (int32_t)lower_blk.size();
            lower_blk.push_back(diag[I]);
            mirror.push_back(diag[I]);
        }
    c->num_lower = (int)lower_blk.size();
    if (100 * E > (int64_t)std::numeric_limits<int32_t>::max())
        throw InvalidArg("mesh too large for int32 staging offsets");
    // Contributions per lower block, in element order within a block (counting sort); each element slot
    // (pairs 0..5, diagonal 6..9) gets its staging offset: records of 10 doubles (9 values + pad, 16-byte
    // aligned) for the off-diagonal blocks, then records of 6 for the diagonal blocks.
    std::vector<int32_t> cptr(c->num_lower + 1, 0);
    std::vector<int32_t> slot_t(10 * E);
    for (int64_t e = 0; e < E; ++e) {
        const int32_t* id = conn + 4 * e;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                int32_t I = std::max(id[a], id[b]), J = std::min(id[a], id[b]);
                const int ps = a * (7 - a) / 2 + (b - a - 1);
                const int32_t t = lower_of_blk[find_blk(I, J)];
                slot_t[10 * e + ps] = t;
                cptr[t + 1]++;
            }
        for (int a = 0; a < 4; ++a) {
            const int32_t t = lower_of_blk[diag[id[a]]];
            slot_t[10 * e + 6 + a] = t;
            cptr[t + 1]++;
        }
    }
    // The element kernel sums, per workgroup of kElemGroup consecutive elements, the contributions its elements
    // make to one lower block (or one node's rhs row) in LDS, in element order, and stores one record per
    // (group, block): records of block t are ordered by group, i.e. by element, so the assembly's sum over them
    // keeps the element order. Per group: its unique blocks q (grp_q), each with the staging offset of its record
    // (q_out) and the LDS sources (element in group * 10 + slot) of its contributions in element order
    // (q_src / src16); the same for the node rows (grp_n, n_out, n_src / nsrc16) with sources element * 4 + corner.
    const int64_t G = (E + kElemGroup - 1) / kElemGroup;
    c->num_groups = G;
    std::fill(cptr.begin(), cptr.end(), 0);
    std::vector<int32_t> rptr(N + 1, 0);
    std::vector<std::array<int32_t, 3>> ctb, ntb;  // (block or node, element in group, source)
    auto group_lists = [&](int64_t g) {
        const int64_t e0 = g * kElemGroup, e1 = std::min<int64_t>(E, e0 + kElemGroup);
        ctb.clear();
        ntb.clear();
        for (int64_t e = e0; e < e1; ++e) {
            const int le = (int)(e - e0);
            for (int sl = 0; sl < 10; ++sl) ctb.push_back({slot_t[10 * e + sl], le, 10 * le + sl});
            for (int q = 0; q < 4; ++q) ntb.push_back({conn[4 * e + q], le, 4 * le + q});
        }
        std::sort(ctb.begin(), ctb.end());
        std::sort(ntb.begin(), ntb.end());
    };
    for (int64_t g = 0; g < G; ++g) {
        group_lists(g);
        for (size_t k = 0; k < ctb.size(); ++k)
            if (k == 0 || ctb[k][0] != ctb[k - 1][0]) cptr[ctb[k][0] + 1]++;
        for (size_t k = 0; k < ntb.size(); ++k)
            if (k == 0 || ntb[k][0] != ntb[k - 1][0]) rptr[ntb[k][0] + 1]++;
    }
    for (int t = 0; t < c->num_lower; ++t) cptr[t + 1] += cptr[t];
    for (int64_t i = 0; i < N; ++i) rptr[i + 1] += rptr[i];
    c->nctr_off = cptr[c->num_lower_off];
    c->stage_len = kPairRec * c->nctr_off + kDiagRec * ((int64_t)cptr[c->num_lower] - c->nctr_off);
    c->rhs_len = kRhsRec * (int64_t)rptr[N];
    if (c->stage_len > (int64_t)std::numeric_limits<int32_t>::max())
        throw InvalidArg("mesh too large for int32 staging offsets");
    std::vector<int32_t> grp_q(G + 1, 0), q_out, q_src(1, 0), grp_n(G + 1, 0), n_out, n_src(1, 0);
    std::vector<uint16_t> src16, nsrc16;
    {
        std::vector<int64_t> pb(cptr.begin(), cptr.end() - 1), pn(rptr.begin(), rptr.end() - 1);
        // A group's records go out longest source list first: one lane sums one record, so lanes of a wave
        // then loop over lists of similar lengths. The sum within a record keeps the element order.
        struct Run {
            size_t first, count;
            int32_t out;
        };
        std::vector<Run> runs;
        auto emit = [&](const std::vector<std::array<int32_t, 3>>& tb, std::vector<int32_t>& out,
                        std::vector<int32_t>& src_ptr, std::vector<uint16_t>& src) {
            std::stable_sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.count > b.count; });
            for (const Run& r : runs) {
                out.push_back(r.out);
                for (size_t k = r.first; k < r.first + r.count; ++k) src.push_back((uint16_t)tb[k][2]);
                src_ptr.push_back((int32_t)src.size());
            }
        };
        for (int64_t g = 0; g < G; ++g) {
            group_lists(g);
            runs.clear();
            for (size_t k = 0; k < ctb.size(); ++k) {
                const int32_t t = ctb[k][0];
                if (k == 0 || t != ctb[k - 1][0]) {
                    const int64_t pos = pb[t]++;
                    runs.push_back({k, 0, (int32_t)(pos < c->nctr_off ? kPairRec * pos
                                                                      : kPairRec * c->nctr_off + kDiagRec * (pos - c->nctr_off))});
                }
                runs.back().count++;
            }
            emit(ctb, q_out, q_src, src16);
            grp_q[g + 1] = (int32_t)q_out.size();
            runs.clear();
            for (size_t k = 0; k < ntb.size(); ++k) {
                const int32_t nd = ntb[k][0];
                if (k == 0 || nd != ntb[k - 1][0]) runs.push_back({k, 0, (int32_t)(kRhsRec * pn[nd]++)});
                runs.back().count++;
            }
            emit(ntb, n_out, n_src, nsrc16);
            grp_n[g + 1] = (int32_t)n_out.size();
        }
    }
    // Node lists of (element, corner) codes in element order, for the adjoint's gather.
    std::vector<int32_t> nptr(N + 1, 0), nlist(4 * E);
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) nptr[conn[4 * e + a] + 1]++;
    for (int64_t i = 0; i < N; ++i) nptr[i + 1] += nptr[i];
    {
        std::vector<int32_t> p(nptr.begin(), nptr.end() - 1);
        for (int64_t e = 0; e < E; ++e)
            for (int a = 0; a < 4; ++a) nlist[p[conn[4 * e + a]]++] = (int32_t)(e << 2) | a;
    }
    hipStream_t s = c->stream;
    c->rowptr_b.alloc(N + 1);
    c->rowptr_b.upload(rp.data(), N + 1, s);
    c->colidx_b.alloc(c->num_blocks);
    c->colidx_b.upload(ci.data(), c->num_blocks, s);
    c->blk_row.alloc(c->num_blocks);
    c->blk_row.upload(blk_row.data(), c->num_blocks, s);
    c->diagblk.alloc(N);
    c->diagblk.upload(diag.data(), N, s);
    c->h_lb_info.resize(c->num_lower);
    c->h_lb_row.resize(c->num_lower);
    c->h_lb_col.resize(c->num_lower);
    for (int t = 0; t < c->num_lower; ++t) {
        c->h_lb_info[t] = make_int4(lower_blk[t], mirror[t], 0, 0);
        c->h_lb_row[t] = blk_row[lower_blk[t]];
        c->h_lb_col[t] = ci[lower_blk[t]];
    }
    c->lb_info.alloc(c->num_lower);
    c->bc_dirty = true;  // ensure_bc fills the boundary bits and uploads lb_info
    c->lower_blk.alloc(c->num_lower);
    c->lower_blk.upload(lower_blk.data(), c->num_lower, s);
    c->mirror_blk.alloc(c->num_lower);
    c->mirror_blk.upload(mirror.data(), c->num_lower, s);
    c->ctr_ptr.alloc(c->num_lower + 1);
    c->ctr_ptr.upload(cptr.data(), c->num_lower + 1, s);
    auto up32 = [&](DevBuf<int>& d, const std::vector<int32_t>& h) {
        d.alloc(std::max<size_t>(1, h.size()));
        d.upload(h.data(), h.size(), s);
    };
    auto up16 = [&](DevBuf<uint16_t>& d, const std::vector<uint16_t>& h) {
        d.alloc(std::max<size_t>(1, h.size()));
        d.upload(h.data(), h.size(), s);
    };
    up32(c->grp_q, grp_q);
    up32(c->q_out, q_out);
    up32(c->q_src, q_src);
    up16(c->src16, src16);
    up32(c->grp_n, grp_n);
    up32(c->n_out, n_out);
    up32(c->n_src, n_src);
    up16(c->nsrc16, nsrc16);
    up32(c->rhs_ptr, rptr);
    c->node_ptr.alloc(N + 1);
    c->node_ptr.upload(nptr.data(), N + 1, s);
    c->node_list.alloc(4 * E);
    c->node_list.upload(nlist.data(), 4 * E, s);
    HIPCHECK(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

void fem_default_options(fem_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->dt = 1e-3;
    o->gravity[0] = 0.0;
    o->gravity[1] = -9.81;
    o->gravity[2] = 0.0;
    o->gravity_enabled = 1;
    o->scheme = FEM_SCHEME_EULER_IMPLICIT;
    o->precision = FEM_PRECISION_FP64;
    o->newton_max_iterations = 1;
    o->cg_tolerance = std::numeric_limits<double>::epsilon();
    o->cg_max_iterations = -1;
    o->inner_tolerance = 1e-4;
    o->max_refinements = 8;
    o->fail_on_nonconvergence = 0;
    o->newton_epsilon = 1e-5;
    o->linear_solver = FEM_LINEAR_SOLVER_CG;
}

int fem_create(int device, fem_ctx** out) {
    if (!out) return FEM_ERR_INVALID;
    *out = nullptr;
    fem_ctx* c = new fem_ctx();
    int rc = guarded(c, [&] {
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw InvalidArg("invalid device ordinal");
        c->device = device;
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        c->scratch = noalloc_scratch_address(device);  // once per device, with the device current
        HIPCHECK(hipHostMalloc((void**)&c->h_sc, 2 * sizeof(CgScalars), hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void**)&c->h_counters, 8 * sizeof(int), hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void**)&c->h_nvar, sizeof(double), hipHostMallocDefault));
        c->nvar.alloc(1);
        c->partial.alloc(4 * kCgBlocks);
        c->scal.alloc(2);
        c->counters.alloc(8);
        HIPCHECK(hipMemset(c->counters.p, 0, 8 * sizeof(int)));
        HIPCHECK(hipMemset(c->scal.p, 0, 2 * sizeof(CgScalars)));
        fem_default_options(&c->opt);
        for (int i = 0; i < 16; ++i) c->predicted_sweep[i] = 32;
        if (const char* e = std::getenv("FEM_CG_KERNELS"))
            c->cg_kernels = std::strcmp(e, "fused") == 0 ? 1 : std::strcmp(e, "split") == 0 ? 2 : 0;
        for (int i = 0; i < 17; ++i) c->predicted_hex[i] = 8;
        std::memset(&c->stats, 0, sizeof(c->stats));
        c->stats.active = 1;
        return FEM_OK;
    });
    if (rc != FEM_OK) {
        fem_destroy(c);
        return rc;
    }
    *out = c;
    return FEM_OK;
}

void fem_destroy(fem_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->h_nvar) (void)hipHostFree(c->h_nvar);
    if (c->h_refine_iters) (void)hipHostFree(c->h_refine_iters);
    if (c->h_bsc) (void)hipHostFree(c->h_bsc);
    if (c->h_live) (void)hipHostFree(c->h_live);
    if (c->h_mcg) (void)hipHostFree(c->h_mcg);
    for (int i = 0; i < 2; ++i) {
        if (c->b_exec[i]) (void)hipGraphExecDestroy(c->b_exec[i]);
        if (c->b_graph[i]) (void)hipGraphDestroy(c->b_graph[i]);
    }
    hex_release(c->hex);
    if (c->lu.ab) (void)hipFree(c->lu.ab);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* fem_last_error(const fem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fem_set_stream(fem_ctx* c, void* s) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (s) {
            if (c->own_stream) HIPCHECK(hipStreamDestroy(c->stream));
            c->stream = (hipStream_t)s;
            c->own_stream = false;
        } else if (!c->own_stream) {
            HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        return FEM_OK;
    });
}

int fem_upload_mesh(fem_ctx* c, int kind, int64_t num_nodes, const double* x0, int64_t num_elements,
                    const int32_t* conn, const double* young, const double* poisson, const double* density) {
    return guarded(c, [&] {
        if (c->mesh_ready) throw InvalidArg("You must set the FemElement type before initializing (mesh already uploaded)");
        if (kind != FEM_ELEMENT_TET_LINEAR && kind != FEM_ELEMENT_TET_COROTATIONAL && kind != FEM_ELEMENT_HEX)
            throw InvalidArg("unknown element kind");
        if (num_nodes <= 0 || num_elements <= 0 || !x0 || !conn || !young || !poisson || !density)
            throw InvalidArg("empty mesh or null pointer");
        if (3 * num_nodes > std::numeric_limits<int32_t>::max() / 2 || num_elements >= (1LL << 27))
            throw InvalidArg("mesh too large for int32 indexing");
        HIPCHECK(hipSetDevice(c->device));
        const int nn = kind == FEM_ELEMENT_HEX ? 8 : 4;
        c->kind = kind;
        c->nn = nn;
        c->num_nodes = num_nodes;
        c->num_elements = num_elements;
        c->ndof = 3 * num_nodes;
        for (int64_t e = 0; e < num_elements; ++e) {
            for (int a = 0; a < nn; ++a) {
                int32_t id = conn[nn * e + a];
                if (id < 0 || id >= num_nodes)
                    throw InvalidArg("Invalid nodeId " + std::to_string(id) + " expected in range [0.." +
                                     std::to_string(num_nodes - 1) + "]");
            }
            check_material(young[e], poisson[e], density[e], e);
            if (kind != FEM_ELEMENT_HEX) {
                // a tet that repeats a node has no volume (Fem3DElementCorotationalTetrahedron.cpp:48-75 rejects
                // it; the linear tet's stiffness would not be finite); the block staging assumes distinct nodes
                const int32_t* id = conn + 4 * e;
                if (id[0] == id[1] || id[0] == id[2] || id[0] == id[3] || id[1] == id[2] || id[1] == id[3] ||
                    id[2] == id[3])
                    throw InvalidArg("element " + std::to_string(e) + " repeats a node");
            }
        }
        c->h_mass_per_node.assign(num_nodes, 0.0);
        c->h_conn.assign(conn, conn + (size_t)nn * num_elements);
        c->h_nu.assign(poisson, poisson + num_elements);
        c->h_rho.assign(density, density + num_elements);
        std::vector<double> mat(4 * num_elements);
        for (int64_t e = 0; e < num_elements; ++e) {
            double E = young[e], nu = poisson[e], rho = density[e];
            mat[4 * e] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
            mat[4 * e + 1] = E / (2.0 * (1 + nu));
            mat[4 * e + 2] = rho;
            mat[4 * e + 3] = E;
        }
        if (kind != FEM_ELEMENT_HEX) {
            for (int64_t e = 0; e < num_elements; ++e) {
                const int32_t* id = conn + 4 * e;
                double vol = tet_volume(x0 + 3 * id[0], x0 + 3 * id[1], x0 + 3 * id[2], x0 + 3 * id[3]);
                if (kind == FEM_ELEMENT_TET_COROTATIONAL && !(std::fabs(6.0 * vol) > 1e-12))
                    throw InvalidArg("Trying to initialize an invalid co-rotational tetrahedron. Matrix V not "
                                     "invertible (element " + std::to_string(e) + ")");
                double mass = vol * density[e];
                for (int a = 0; a < 4; ++a) c->h_mass_per_node[id[a]] += mass / 4;
            }
            {
                // diag(M): consistent, the elements' diagonal entries 2 rho V / 20 in element order; lumped, the
                // row sums rho V / 4 (the node masses)
                std::vector<double> dc((size_t)3 * num_nodes, 0.0), dl((size_t)3 * num_nodes, 0.0);
                for (int64_t e = 0; e < num_elements; ++e) {
                    const int32_t* id = conn + 4 * e;
                    const double m = tet_volume(x0 + 3 * id[0], x0 + 3 * id[1], x0 + 3 * id[2], x0 + 3 * id[3]) *
                                     density[e] / 20.0;
                    for (int a = 0; a < 4; ++a)
                        for (int k = 0; k < 3; ++k) dc[3 * (size_t)id[a] + k] += 2.0 * m;
                }
                for (int64_t i = 0; i < num_nodes; ++i)
                    for (int k = 0; k < 3; ++k) dl[3 * (size_t)i + k] = c->h_mass_per_node[i];
                c->mdiag_cons.alloc(dc.size());
                c->mdiag_cons.upload(dc.data(), dc.size(), c->stream);
                c->mdiag_lump.alloc(dl.size());
                c->mdiag_lump.upload(dl.data(), dl.size(), c->stream);
                HIPCHECK(hipStreamSynchronize(c->stream));  // the host tables are freed on return
            }
            build_tet_maps(c, conn);
            std::vector<int4> c4(num_elements);
            for (int64_t e = 0; e < num_elements; ++e)
                c4[e] = make_int4(conn[4 * e], conn[4 * e + 1], conn[4 * e + 2], conn[4 * e + 3]);
            c->conn4.alloc(num_elements);
            c->conn4.upload(c4.data(), num_elements, c->stream);
            c->rhs_rec.alloc((size_t)c->rhs_len);
        } else {
            hex_setup(c->hex, num_nodes, x0, num_elements, conn, mat.data(), c->h_mass_per_node, c->h_rowptr_b,
                      c->h_colidx_b, c->stream);
            c->num_blocks = (int64_t)c->h_colidx_b.size();
        }
        c->mat.alloc(num_elements);
        c->mat.upload(reinterpret_cast<const double4*>(mat.data()), num_elements, c->stream);
        size_t n = (size_t)c->ndof;
        c->x0.alloc(n);
        c->x0.upload(x0, n, c->stream);
        c->x.alloc(n);
        c->x.upload(x0, n, c->stream);
        c->v.alloc(n);
        HIPCHECK(hipMemsetAsync(c->v.p, 0, n * sizeof(double), c->stream));
        c->xs.alloc(n);
        c->vs.alloc(n);
        c->fg.alloc(n);
        c->fext.alloc(n);
        c->b.alloc(n);
        c->dv.alloc(n);
        c->dvprev.alloc(n);
        c->r64.alloc(n);
        c->bcmask.alloc(n);
        c->h_bcmask.assign(n, 0);
        c->h_bc_dofs.clear();
        c->bc_dirty = true;
        c->fg_dirty = true;
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->mesh_ready = true;
        c->ode_init = 0;  // a new representation: no F, M, D, K yet
        c->m_lumped = false;
        c->last_precision = -1;
        alloc_precision_buffers(c);
        return FEM_OK;
    });
}

int fem_add_boundary_condition_dof(fem_ctx* c, int64_t dof) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("Number of dof per node = 0. Make sure to call setNumDof() first");
        if (dof < 0 || dof >= c->ndof) throw InvalidArg("Invalid dof " + std::to_string(dof));
        if (!c->h_bcmask[dof]) {
            c->h_bcmask[dof] = 1;
            c->h_bc_dofs.push_back(dof);
            c->bc_dirty = true;
            c->linear_matrix_ready = false;
        }
        return FEM_OK;
    });
}

int fem_add_boundary_condition_node(fem_ctx* c, int64_t node) {
    if (!c) return FEM_ERR_INVALID;
    if (c->mesh_ready && (node < 0 || node >= c->num_nodes))
        return fail(c, FEM_ERR_INVALID, "Invalid nodeId " + std::to_string(node));
    for (int k = 0; k < 3; ++k) {
        int rc = fem_add_boundary_condition_dof(c, 3 * node + k);
        if (rc) return rc;
    }
    return FEM_OK;
}

int fem_set_options(fem_ctx* c, const fem_options* o) {
    return guarded(c, [&] {
        if (!o) throw InvalidArg("null options");
        if (!(o->dt > 0.0)) throw InvalidArg("dt must be positive");
        if (o->scheme < FEM_SCHEME_EULER_IMPLICIT || o->scheme > FEM_SCHEME_LINEAR_RUNGE_KUTTA_4)
            throw InvalidArg("unknown integration scheme");
        if (!std::isfinite(o->rayleigh_mass) || !std::isfinite(o->rayleigh_stiffness))
            throw InvalidArg("Rayleigh damping coefficients must be finite");
        if (o->precision != FEM_PRECISION_FP64 && o->precision != FEM_PRECISION_FP32_REFINED)
            throw InvalidArg("unknown precision");
        if (o->precision == FEM_PRECISION_FP32_REFINED && o->max_refinements < 1)
            throw InvalidArg("FP32 precision needs at least one refinement sweep (max_refinements >= 1)");
        if (o->newton_max_iterations < 1) throw InvalidArg("The maximum iteration needs to be at least 1");
        if (!(o->newton_epsilon >= 0.0)) throw InvalidArg("The epsilon convergence cannot be negative");
        if (o->linear_solver != FEM_LINEAR_SOLVER_LU && o->linear_solver != FEM_LINEAR_SOLVER_CG)
            throw InvalidArg("unknown linear solver");
        if (o->linear_solver == FEM_LINEAR_SOLVER_LU && c->mesh_ready && (c->kind == FEM_ELEMENT_HEX || c->num_instances > 0))
            throw InvalidArg("the direct solver needs the assembled operator of a tetrahedral mesh of one instance");
        if (!(o->cg_tolerance >= 0.0)) throw InvalidArg("negative tolerance");
        bool grav = o->gravity_enabled != c->opt.gravity_enabled || o->gravity[0] != c->opt.gravity[0] ||
                    o->gravity[1] != c->opt.gravity[1] || o->gravity[2] != c->opt.gravity[2];
        if (o->scheme != c->opt.scheme || o->dt != c->opt.dt || o->rayleigh_mass != c->opt.rayleigh_mass ||
            o->rayleigh_stiffness != c->opt.rayleigh_stiffness || o->mass_lumping != c->opt.mass_lumping)
            c->linear_matrix_ready = false;
        c->opt = *o;
        if (grav) c->fg_dirty = true;
        if (c->mesh_ready) {
            HIPCHECK(hipSetDevice(c->device));
            alloc_precision_buffers(c);
        }
        return FEM_OK;
    });
}

int fem_set_external_force(fem_ctx* c, const double* f) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        HIPCHECK(hipSetDevice(c->device));
        if (f) {
            c->fext.upload(f, (size_t)c->ndof, c->stream);
            HIPCHECK(hipStreamSynchronize(c->stream));
            c->have_fext = true;
        } else {
            c->have_fext = false;
        }
        return FEM_OK;
    });
}

int fem_add_external_generalized_force(fem_ctx* c, int64_t element, const double* coordinate, const double* force,
                                       const double* stiffness, const double* damping) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        if (!coordinate || !force) throw InvalidArg("Invalid localization (null coordinate or force)");
        if (element < 0 || element >= c->num_elements) throw InvalidArg("Invalid localization (element out of range)");
        if ((stiffness || damping) && scheme_is_linear(c->opt.scheme))
            throw InvalidArg("external stiffness and damping are not supported by the linear (constant-operator) schemes");
        const int nn = c->nn;
        if (c->h_loc_f.size() != (size_t)c->ndof) c->h_loc_f.assign((size_t)c->ndof, 0.0);
        const int32_t* ids = c->h_conn.data() + (size_t)nn * element;
        for (int a = 0; a < nn; ++a)
            for (int k = 0; k < 3; ++k) c->h_loc_f[3 * (size_t)ids[a] + k] += force[k] * coordinate[a];
        if (stiffness || damping)
            for (int a = 0; a < nn; ++a)
                for (int b = 0; b < nn; ++b) {
                    auto& blk = c->h_loc_blk[std::make_pair((int64_t)ids[a], (int64_t)ids[b])];
                    const double w = coordinate[a] * coordinate[b];
                    for (int k = 0; k < 9; ++k) {
                        if (stiffness) blk[k] += w * stiffness[k];
                        if (damping) blk[9 + k] += w * damping[k];
                    }
                }
        c->loc_pending = true;
        return FEM_OK;
    });
}

namespace {
// One compliance column through the step's own solver: b with its boundary rows zeroed as the rhs, the solution
// with its boundary rows zeroed.
void compliance_column(fem_ctx* c, const double* b, double* out) {
    const size_t n = (size_t)c->ndof;
    launch_masked_copy((int64_t)n, b, c->bcmask.p, c->b.p, c->stream);
    if (c->kind == FEM_ELEMENT_HEX) {
        hex_solve(c->hex, hex_args(c), c->stats, c->stream);
    } else if (c->num_instances > 0) {
        HIPCHECK(hipMemsetAsync(c->dv.p, 0, n * sizeof(double), c->stream));
        batch_linear_solve(c);
    } else {
        linear_solve(c);
    }
    launch_masked_copy((int64_t)n, c->dv.p, c->bcmask.p, out, c->stream);
}

// Compliance columns of a tetrahedral mesh, up to 8 at a time through the multi-column FP64 CG: the columns
// (device, column-major) are interleaved [dof][kw] with their boundary rows zeroed, solved together against
// S64, and scattered back with the boundary rows of the solutions zeroed. In the FP32 + refinement mode a
// remainder of fewer than 5 columns goes column by column through the step's FP32 solver instead: on the
// 40^3 block a fixed-point constraint's 3 columns take 2.04 ms that way against 2.19 ms in one 4-column FP64
// pass, while a full 8-column pass costs 0.35 ms per column.
// Largest DOF count whose multi-column CG buffers fit the allocation-free workspace.
int64_t noalloc_max_dof() {
    size_t n = kNoAllocScratchBytes / (49 * sizeof(double));
    while (n > 0 && mcg_layout(n).total > kNoAllocScratchBytes) --n;
    return (int64_t)n;
}

// Work buffers of the multi-column CG: 8 interleaved columns of n DOFs for b, x, r, z, p, q; n for the
// Jacobi inverse; 3 x kCgBlocks x 8 block sums; 8 column states. status: host copy of the column states.
struct McgWork {
    double *b, *x, *r, *z, *p, *q, *dinv, *partial;
    McgColumn* col;
    McgColumn* status;
};

// Columns [0, ncols) of b (device, column-major) through the multi-column FP64 CG, up to 8 at a time: each
// batch is interleaved [dof][kw] with its boundary rows zeroed, solved against S64, and scattered back with
// the boundary rows of the solutions zeroed. The host reads the column states between chunks of launches
// (iterations past a column's finish leave it unchanged, so the launch count does not change the result).
void mcg_columns(fem_ctx* c, const double* b, int64_t ncols, double* out, const McgWork& w) {
    const size_t n = (size_t)c->ndof;
    const int max_it = (int)std::min<int64_t>(max_iters(c), std::numeric_limits<int>::max() - 1);
    for (int64_t done = 0; done < ncols;) {
        const int64_t left = ncols - done;
        const int kw = left >= 5 ? 8 : left >= 3 ? 4 : (int)left;
        const int cnt = (int)std::min<int64_t>(left, kw);
        launch_mcg_gather((int64_t)n, kw, cnt, b + (size_t)done * n, c->bcmask.p, w.b, c->stream);
        McgArgs a;
        a.kw = kw;
        a.num_nodes = (int)c->num_nodes;
        a.rowptr = c->rowptr_b.p;
        a.colidx = c->colidx_b.p;
        a.diagblk = c->diagblk.p;
        a.S = c->S64.p;
        a.b = w.b;
        a.x = w.x;
        a.r = w.r;
        a.z = w.z;
        a.p = w.p;
        a.q = w.q;
        a.dinv = w.dinv;
        a.partial = w.partial;
        a.col = w.col;
        a.tol = c->opt.cg_tolerance;
        a.max_iter = max_it;
        launch_mcg_init(a, c->stream);
        const int pred = std::max(4, c->predicted_iters + 2);
        int k = 0;
        for (; k < std::min(pred, max_it); ++k) launch_mcg_iteration(a, k, c->stream);
        for (;;) {
            HIPCHECK(hipMemcpyAsync(w.status, w.col, kw * sizeof(McgColumn), hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
            bool running = false;
            for (int j = 0; j < kw; ++j) running = running || w.status[j].fin == kMcgRunning;
            if (!running || k >= max_it) break;
            const int chunk = std::max(2, pred / 8);
            for (int i = 0; i < chunk && k < max_it; ++i, ++k) launch_mcg_iteration(a, k, c->stream);
        }
        launch_mcg_scatter((int64_t)n, kw, cnt, w.x, c->bcmask.p, out + (size_t)done * n, c->stream);
        done += cnt;
    }
}

// Compliance columns through the band LU (LinearSparseSolveAndInverseLU::solve): each column of b with its
// boundary rows zeroed into tmp (tmp_cols columns of ndof at a time), solved into out, boundary rows of the
// solutions zeroed (DeformableRepresentation::applyCompliance, Physics/DeformableRepresentation.cpp:175-190).
void compliance_lu(fem_ctx* c, const double* b, int64_t ncols, double* out, double* tmp, int64_t tmp_cols) {
    const int64_t n = c->ndof;
    for (int64_t j0 = 0; j0 < ncols; j0 += tmp_cols) {
        const int cnt = (int)std::min<int64_t>(tmp_cols, ncols - j0);
        for (int j = 0; j < cnt; ++j)
            launch_masked_copy(n, b + (size_t)(j0 + j) * n, c->bcmask.p, tmp + (size_t)j * n, c->stream);
        launch_band_solve(c->lu, cnt, tmp, n, out + (size_t)j0 * n, n, c->stream);
        launch_zero_masked_rows(n, cnt, c->bcmask.p, out + (size_t)j0 * n, c->stream);
    }
}

// Compliance columns of a tetrahedral mesh through the multi-column FP64 CG (mcg_columns) on the context's
// own work buffers. In the FP32 + refinement mode a remainder of fewer than 5 columns goes column by column
// through the step's FP32 solver instead: on the 40^3 block a fixed-point constraint's 3 columns take 2.04 ms
// that way against 2.19 ms in one 4-column FP64 pass, while a full 8-column pass costs 0.35 ms per column.
void compliance_tet(fem_ctx* c, const double* b, int64_t ncols, double* out) {
    const size_t n = (size_t)c->ndof;
    if (c->opt.linear_solver == FEM_LINEAR_SOLVER_LU) {
        ensure_lu(c);
        const int64_t tc = std::max<int64_t>(1, std::min<int64_t>(ncols, 64));
        DevBuf<double> tmp;
        tmp.alloc(n * (size_t)tc);
        compliance_lu(c, b, ncols, out, tmp.p, tc);
        return;
    }
    if (c->opt.precision == FEM_PRECISION_FP32_REFINED) {
        const int64_t rem = ncols % 8, full = ncols - rem;
        if (rem > 0 && rem < 5) {
            for (int64_t j = full; j < ncols; ++j) compliance_column(c, b + (size_t)j * n, out + (size_t)j * n);
            ncols = full;
        }
    }
    if (c->mcg_b.n != n * 8) {
        for (DevBuf<double>* v : {&c->mcg_b, &c->mcg_x, &c->mcg_r, &c->mcg_z, &c->mcg_p, &c->mcg_q}) v->alloc(n * 8);
        c->mcg_dinv.alloc(n);
        c->mcg_part.alloc((size_t)3 * kCgBlocks * 8);
        c->mcg_col.alloc(8);
        if (!c->h_mcg) HIPCHECK(hipHostMalloc((void**)&c->h_mcg, 8 * sizeof(McgColumn), hipHostMallocDefault));
    }
    const McgWork w{c->mcg_b.p, c->mcg_x.p, c->mcg_r.p, c->mcg_z.p, c->mcg_p.p, c->mcg_q.p,
                    c->mcg_dinv.p, c->mcg_part.p, c->mcg_col.p, c->h_mcg};
    mcg_columns(c, b, ncols, out, w);
}

// C b for num_columns device columns (column-major) on the context stream.
void compliance_device(fem_ctx* c, const double* b, int64_t num_columns, double* out) {
    ensure_bc(c);
    alloc_precision_buffers(c);
    const size_t n = (size_t)c->ndof;
    // the step's statistics and iteration predictions are the step's, not the compliance solves'
    const fem_stats saved = c->stats;
    const int saved_iters = c->predicted_iters, saved_sweeps = c->predicted_sweeps;
    int saved_sweep[16], saved_hex[17], saved_batch[17];
    std::memcpy(saved_sweep, c->predicted_sweep, sizeof saved_sweep);
    std::memcpy(saved_hex, c->predicted_hex, sizeof saved_hex);
    std::memcpy(saved_batch, c->b_predicted, sizeof saved_batch);
    if (c->kind != FEM_ELEMENT_HEX && c->num_instances == 0) {
        compliance_tet(c, b, num_columns, out);
    } else {
        for (int64_t j = 0; j < num_columns; ++j) compliance_column(c, b + (size_t)j * n, out + (size_t)j * n);
    }
    c->stats = saved;
    c->predicted_iters = saved_iters;
    c->predicted_sweeps = saved_sweeps;
    std::memcpy(c->predicted_sweep, saved_sweep, sizeof saved_sweep);
    std::memcpy(c->predicted_hex, saved_hex, sizeof saved_hex);
    std::memcpy(c->b_predicted, saved_batch, sizeof saved_batch);
}
}  // namespace

int fem_apply_compliance(fem_ctx* c, const double* b, int64_t num_columns, double* out) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        if (num_columns < 0 || (num_columns > 0 && (!b || !out))) throw InvalidArg("null or negative columns");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        HIPCHECK(hipSetDevice(c->device));
        const size_t n = (size_t)c->ndof;
        // stage through device memory, up to 64 columns at a time
        const int64_t chunk = std::min<int64_t>(std::max<int64_t>(num_columns, 1), 64);
        DevBuf<double> db, dx;
        db.alloc(n * chunk);
        dx.alloc(n * chunk);
        for (int64_t j0 = 0; j0 < num_columns; j0 += chunk) {
            const int64_t m = std::min(chunk, num_columns - j0);
            db.upload(b + (size_t)j0 * n, n * m, c->stream);
            compliance_device(c, db.p, m, dx.p);
            HIPCHECK(hipMemcpyAsync(out + (size_t)j0 * n, dx.p, n * m * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
        return FEM_OK;
    });
}

int fem_apply_compliance_device(fem_ctx* c, const double* b, int64_t num_columns, double* out) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        if (num_columns < 0 || (num_columns > 0 && (!b || !out))) throw InvalidArg("null or negative columns");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        HIPCHECK(hipSetDevice(c->device));
        compliance_device(c, b, num_columns, out);
        return FEM_OK;
    });
}

int fem_apply_compliance_device_noalloc(fem_ctx* c, const double* b, int64_t num_columns, double* out) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        if (num_columns < 0 || (num_columns > 0 && (!b || !out))) throw InvalidArg("null or negative columns");
        if (c->kind == FEM_ELEMENT_HEX || c->num_instances > 0)
            throw InvalidArg("the allocation-free compliance needs a tetrahedral mesh of one instance");
        const McgLayout L = mcg_layout((size_t)c->ndof);
        if (L.total > kNoAllocScratchBytes)
            throw InvalidArg("the allocation-free compliance supports at most " +
                             std::to_string(noalloc_max_dof()) + " DOFs (" +
                             std::to_string(kNoAllocScratchBytes >> 20) + " MiB workspace); this mesh has " +
                             std::to_string(c->ndof));
        if (!c->scratch) throw InvalidArg("no workspace on this device");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        if (c->bc_dirty)
            return fail(c, FEM_ERR_ORDER, "boundary conditions changed since the last step: step first");
        HIPCHECK(hipSetDevice(c->device));
        char* base = (char*)c->scratch;
        if (c->opt.linear_solver == FEM_LINEAR_SOLVER_LU) {
            ensure_lu(c, false);  // refactoring reuses the band storage of the last step's factorization
            const int64_t tc = (int64_t)(kNoAllocScratchBytes / ((size_t)c->ndof * sizeof(double)));
            compliance_lu(c, b, num_columns, out, (double*)base, tc);
            return FEM_OK;
        }
        McgColumn status[8];
        const McgWork w{(double*)(base + L.b), (double*)(base + L.x), (double*)(base + L.r),
                        (double*)(base + L.z), (double*)(base + L.p), (double*)(base + L.q),
                        (double*)(base + L.dinv), (double*)(base + L.partial), (McgColumn*)(base + L.col), status};
        mcg_columns(c, b, num_columns, out, w);
        return FEM_OK;
    });
}

namespace {
// Columns [j0, j0 + cnt) of the dense compliance into out (ndof x cnt, column-major): identity columns solved
// through the band LU of the operator, then the boundary rows and columns zeroed.
void compliance_columns_dense(fem_ctx* c, int64_t j0, int cnt, double* out) {
    const int64_t n = c->ndof;
    launch_identity_columns(n, j0, cnt, out, c->stream);
    launch_band_solve(c->lu, cnt, out, n, out, n, c->stream, j0);
    launch_zero_bc_rows_cols(n, j0, cnt, c->bcmask.p, out, c->stream);
}

// OdeSolverLinearEulerImplicit forms the compliance on its first pass and keeps it (computeCompliance = true,
// Math/OdeSolverLinearEulerImplicit.cpp:43-51; "the compliance matrix is constant and used in all following
// calls"). Kept here when it fits kDenseKeepBytes; the later steps then apply it (lu_linear_solve) and
// fem_get_compliance_matrix returns it.
constexpr size_t kDenseKeepBytes = size_t(1) << 30;

void keep_dense_compliance(fem_ctx* c) {
    const int64_t n = c->ndof;
    if ((size_t)n * (size_t)n * sizeof(double) > kDenseKeepBytes) return;
    ensure_bc(c);
    ensure_lu(c);
    if (c->dense_c.n != (size_t)(n * n)) c->dense_c.alloc((size_t)(n * n));
    if (c->gemv_part.n != (size_t)(gemv_chunks(n) * n)) c->gemv_part.alloc((size_t)(gemv_chunks(n) * n));
    compliance_columns_dense(c, 0, (int)n, c->dense_c.p);
    c->dense_c_valid = true;
}

void check_dense_compliance(fem_ctx* c) {
    if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
    if (c->kind == FEM_ELEMENT_HEX || c->num_instances > 0)
        throw InvalidArg("the dense compliance needs the assembled operator of a tetrahedral mesh of one instance");
}
}  // namespace

int fem_get_compliance_matrix(fem_ctx* c, double* out) {
    return guarded(c, [&] {
        check_dense_compliance(c);
        if (!out) throw InvalidArg("null output");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        HIPCHECK(hipSetDevice(c->device));
        ensure_bc(c);
        const int64_t n = c->ndof;
        if (c->dense_c_valid) {  // the kept compliance of the current factors
            HIPCHECK(hipMemcpyAsync(out, c->dense_c.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
            return FEM_OK;
        }
        ensure_lu(c);
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)((size_t(256) << 20) / (8 * n))));
        DevBuf<double> buf;
        buf.alloc((size_t)n * chunk);
        for (int64_t j0 = 0; j0 < n; j0 += chunk) {
            const int cnt = (int)std::min(chunk, n - j0);
            compliance_columns_dense(c, j0, cnt, buf.p);
            HIPCHECK(hipMemcpyAsync(out + (size_t)j0 * n, buf.p, (size_t)n * cnt * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
        return FEM_OK;
    });
}

int fem_get_compliance_matrix_device(fem_ctx* c, double* out) {
    return guarded(c, [&] {
        check_dense_compliance(c);
        if (!out) throw InvalidArg("null output");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        HIPCHECK(hipSetDevice(c->device));
        ensure_bc(c);
        const int64_t n = c->ndof;
        if (c->dense_c_valid) {  // the kept compliance of the current factors
            HIPCHECK(hipMemcpyAsync(out, c->dense_c.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice,
                                    c->stream));
            return FEM_OK;
        }
        ensure_lu(c);
        for (int64_t j0 = 0; j0 < n; j0 += 65536)
            compliance_columns_dense(c, j0, (int)std::min<int64_t>(65536, n - j0), out + (size_t)j0 * n);
        return FEM_OK;
    });
}

int fem_constraint_compliance(fem_ctx* c, int64_t element, const double* coordinate, int num_directions,
                              const double* directions, double scale, double* cht, double* hcht) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        if (!coordinate || !directions) throw InvalidArg("Invalid localization (null coordinate or directions)");
        if (element < 0 || element >= c->num_elements) throw InvalidArg("Invalid localization (element out of range)");
        if (num_directions < 1 || num_directions > 8) throw InvalidArg("num_directions must be 1 to 8");
        if (!c->have_operator) return fail(c, FEM_ERR_ORDER, "no system matrix yet: the solver has not stepped");
        HIPCHECK(hipSetDevice(c->device));
        const int nn = c->nn, m = num_directions;
        const size_t n = (size_t)c->ndof;
        // The non-zeros of one row of H in DOF order, as the constraint's SparseVector holds them: DOF
        // 3 node_a + k for every node with a non-zero coordinate.
        const int32_t* ids = c->h_conn.data() + (size_t)nn * element;
        std::vector<std::pair<int64_t, int>> nz;  // (DOF, 3 a + k)
        for (int a = 0; a < nn; ++a)
            if (coordinate[a] != 0.0)
                for (int k = 0; k < 3; ++k) nz.emplace_back(3 * (int64_t)ids[a] + k, 3 * a + k);
        std::sort(nz.begin(), nz.end());
        const int cnt = (int)nz.size();
        std::vector<double> val((size_t)m * cnt);  // H_j at its non-zeros
        std::vector<int64_t> idx((size_t)m * cnt);  // the same positions in each of the m columns
        for (int j = 0; j < m; ++j)
            for (int t = 0; t < cnt; ++t) {
                const int a = nz[t].second / 3, k = nz[t].second % 3;
                val[(size_t)j * cnt + t] = coordinate[a] * directions[3 * j + k] * scale;
                idx[(size_t)j * cnt + t] = (int64_t)j * (int64_t)n + nz[t].first;
            }
        if (c->con_b.n != n * 8) {
            c->con_b.alloc(n * 8);
            c->con_x.alloc(n * 8);
            c->con_idx.alloc(8 * 8 * 3);
            c->con_val.alloc(8 * 8 * 3);
            c->con_out.alloc(8 * 8 * 3);
        }
        HIPCHECK(hipMemsetAsync(c->con_b.p, 0, n * m * sizeof(double), c->stream));
        if (cnt > 0) {
            HIPCHECK(hipMemcpyAsync(c->con_idx.p, idx.data(), idx.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                    c->stream));
            HIPCHECK(hipMemcpyAsync(c->con_val.p, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice,
                                    c->stream));
            launch_scatter_values(m * cnt, c->con_idx.p, c->con_val.p, c->con_b.p, c->stream);
        }
        compliance_device(c, c->con_b.p, m, c->con_x.p);  // C H^T, boundary rows zero
        std::vector<double> g((size_t)m * cnt);
        if (hcht && cnt > 0) {
            launch_gather_values(m * cnt, c->con_idx.p, c->con_x.p, c->con_out.p, c->stream);
            HIPCHECK(hipMemcpyAsync(g.data(), c->con_out.p, g.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
        }
        if (cht)
            HIPCHECK(hipMemcpyAsync(cht, c->con_x.p, n * m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (hcht)
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) {
                    double s = 0.0;  // H_i (C H_j^T) over H_i's non-zeros in DOF order
                    for (int t = 0; t < cnt; ++t) s += val[(size_t)i * cnt + t] * g[(size_t)j * cnt + t];
                    hcht[i * m + j] = s;
                }
        return FEM_OK;
    });
}

int fem_set_state(fem_ctx* c, const double* x, const double* v) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        HIPCHECK(hipSetDevice(c->device));
        if (x) c->x.upload(x, (size_t)c->ndof, c->stream);
        if (v) c->v.upload(v, (size_t)c->ndof, c->stream);
        if (v) c->v_static_zero = false;
        HIPCHECK(hipStreamSynchronize(c->stream));
        return FEM_OK;
    });
}

int fem_get_state(fem_ctx* c, double* x, double* v) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        HIPCHECK(hipSetDevice(c->device));
        size_t bytes = (size_t)c->ndof * sizeof(double);
        if (x) HIPCHECK(hipMemcpyAsync(x, c->x.p, bytes, hipMemcpyDeviceToHost, c->stream));
        if (v) HIPCHECK(hipMemcpyAsync(v, c->v.p, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        return FEM_OK;
    });
}

int fem_set_state_device(fem_ctx* c, const double* x, const double* v) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        HIPCHECK(hipSetDevice(c->device));
        size_t bytes = (size_t)c->ndof * sizeof(double);
        if (x) HIPCHECK(hipMemcpyAsync(c->x.p, x, bytes, hipMemcpyDeviceToDevice, c->stream));
        if (v) HIPCHECK(hipMemcpyAsync(c->v.p, v, bytes, hipMemcpyDeviceToDevice, c->stream));
        if (v) c->v_static_zero = false;
        return FEM_OK;
    });
}

int fem_get_state_device(fem_ctx* c, double* x, double* v) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        HIPCHECK(hipSetDevice(c->device));
        size_t bytes = (size_t)c->ndof * sizeof(double);
        if (x) HIPCHECK(hipMemcpyAsync(x, c->x.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        if (v) HIPCHECK(hipMemcpyAsync(v, c->v.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        return FEM_OK;
    });
}

namespace {

int step_checks(fem_ctx* c, int32_t nsteps) {
    if (!c->mesh_ready) return fail(c, FEM_ERR_ORDER, "Initial state has not been set yet");
    if (nsteps < 0) throw InvalidArg("negative step count");
    if (!c->stats.active) return FEM_INACTIVE;
    HIPCHECK(hipSetDevice(c->device));
    ensure_bc(c);
    alloc_precision_buffers(c);
    if (c->opt.linear_solver == FEM_LINEAR_SOLVER_LU && (c->kind == FEM_ELEMENT_HEX || c->num_instances > 0))
        throw InvalidArg("the direct solver needs the assembled operator of a tetrahedral mesh of one instance");
    if (c->opt.scheme == FEM_SCHEME_LINEAR_EULER_IMPLICIT && c->kind == FEM_ELEMENT_TET_COROTATIONAL &&
        c->opt.rayleigh_stiffness != 0.0)
        throw InvalidArg("Rayleigh stiffness damping with the linear implicit scheme needs linear elements");
    // Once K exists (an implicit or static step formed it), computeF takes the Rayleigh stiffness force from that
    // global K (addRayleighDampingForce with useGlobalStiffnessMatrix = hasK(), Physics/FemRepresentation.cpp:
    // 596-612), which the explicit and Runge-Kutta schemes never reform and the static scheme reforms only after
    // F (Math/OdeEquation.cpp:78-103): stale for corotational elements and holding the localized stiffness of the
    // step that formed it, neither restated here (the current element K is used). A static step after a static
    // one has v = 0 and no such force.
    const int mode = scheme_mode(c->opt.scheme);
    if (mode != 0 && c->opt.rayleigh_stiffness != 0.0 && (c->ode_init & 8) != 0 &&
        (c->kind == FEM_ELEMENT_TET_COROTATIONAL || c->k_has_ext) && !(mode == 2 && c->v_static_zero))
        throw InvalidArg("Rayleigh stiffness damping in an explicit, Runge-Kutta or static step after a step whose "
                         "global K differs from the current element K (corotational elements, or a localized "
                         "stiffness; the reference's stale global K) is not supported");
    return FEM_OK;
}

// State after step i of a call into the history (device, rows of ndof), i = 0 for the call's start.
void record_state(fem_ctx* c, double* x_hist, double* v_hist, int i) {
    if (!x_hist) return;
    const size_t n = (size_t)c->ndof;
    HIPCHECK(hipMemcpyAsync(x_hist + i * n, c->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(v_hist + i * n, c->v.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

// Element flag counts and the invalid-state flag of the call (read into h_counters by the caller).
int finish_steps(fem_ctx* c) {
    c->timer.resolve();
    c->timer.enabled = false;
    c->stats.det_r_invalid = c->h_counters[0];
    c->stats.g_singular = c->h_counters[1];
    c->stats.fp64_polar_fallbacks = c->h_counters[2];
    if (c->h_counters[3]) {
        c->stats.active = 0;
        return FEM_INACTIVE;
    }
    return FEM_OK;
}

// The synchronous step loop: the host reads each FP64 solve's outcome (each FP32 refinement's per step) before
// the next one goes out.
int step_loop(fem_ctx* c, int32_t nsteps, double* x_hist, double* v_hist) {
    record_state(c, x_hist, v_hist, 0);
    for (int i = 0; i < nsteps; ++i) {
        c->timer.enabled = c->timer_every > 0 && (c->timer_count++ % c->timer_every) == 0;
        prepare_external(c);
        step_once(c);
        c->v_static_zero = scheme_mode(c->opt.scheme) == 2;  // OdeSolverStatic zeroes the velocities
        clear_external(c);
        record_state(c, x_hist, v_hist, i + 1);
        if (c->opt.fail_on_nonconvergence && !c->stats.converged)
            return fail(c, FEM_ERR_NOT_CONVERGED, "linear solve did not converge");
    }
    HIPCHECK(hipMemcpyAsync(c->h_counters, c->counters.p, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return finish_steps(c);
}

}  // namespace

int fem_step(fem_ctx* c, int32_t nsteps) {
    return guarded(c, [&] {
        const int rc = step_checks(c, nsteps);
        if (rc != FEM_OK) return rc;
        return step_loop(c, nsteps, nullptr, nullptr);
    });
}

int fem_step_record(fem_ctx* c, int32_t nsteps, double* x_hist, double* v_hist) {
    return guarded(c, [&] {
        if (!x_hist || !v_hist) throw InvalidArg("null state history");
        const int rc = step_checks(c, nsteps);
        if (rc != FEM_OK) return rc;
        return step_loop(c, nsteps, x_hist, v_hist);
    });
}

int fem_synchronize(fem_ctx* c) {
    return guarded(c, [&] {
        HIPCHECK(hipStreamSynchronize(c->stream));
        return FEM_OK;
    });
}

int fem_query(fem_ctx* c, fem_stats* s) {
    if (!c || !s) return FEM_ERR_INVALID;
    *s = c->stats;
    return FEM_OK;
}

int fem_enable_phase_timers(fem_ctx* c, int enable) {
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->timer.resolve();
        c->timer.reset();
        c->timer_every = enable > 0 ? enable : 0;
        c->timer_count = 0;
        c->timer.enabled = false;
        return FEM_OK;
    });
}

int fem_time_operator(fem_ctx* c, int reps, double* avg_ms) {
    return guarded(c, [&] {
        if (!avg_ms || reps < 1) throw InvalidArg("reps must be positive and avg_ms non-null");
        if (c->stats.steps < 1) return fail(c, FEM_ERR_ORDER, "step at least once before timing the operator");
        HIPCHECK(hipSetDevice(c->device));
        const bool fp32 = c->opt.precision == FEM_PRECISION_FP32_REFINED;
        if (c->kind == FEM_ELEMENT_HEX) {
            *avg_ms = hex_time_operator(c->hex, hex_args(c), fp32, reps, c->stream);
            return FEM_OK;
        }
        if (c->num_instances > 0) throw InvalidArg("operator replay timing covers single meshes");
        if (!c->time_sc.p) c->time_sc.alloc(1);
        // never converges (threshold < 0), fixed rz so every replay does identical work on CG scratch
        CgScalars init{};
        init.rz[0] = init.rz[1] = 1.0;
        init.bb = 1.0;
        init.threshold = -1.0;
        HIPCHECK(hipMemcpyAsync(c->time_sc.p, &init, sizeof(init), hipMemcpyHostToDevice, c->stream));
        CgArgs a = cg_args(c, fp32, 0.0);
        a.sc = c->time_sc.p;
        const int big = std::numeric_limits<int>::max() - 1;
        // the dominant kernel of an iteration: the fused kernel, or the operator kernel of the split pair
        const bool fused = use_fused_cg(c, fp32);
        auto launch = [&] {
            if (fused) launch_cg_fused(a, 1, big, c->stream);
            else launch_cg_apply(a, 1, big, c->stream);
        };
        launch();  // warm-up
        // each launch carries its own start/stop events in its dispatch packet (the kernel's execution, as a
        // profiler's kernel trace reports it, without the gaps between back-to-back dispatches)
        ScopedEvents events(2 * (size_t)reps);
        const std::vector<hipEvent_t>& ev = events.ev;
        for (int r = 0; r < reps; ++r) {
            g_launch_timing.start = ev[2 * r];
            g_launch_timing.stop = ev[2 * r + 1];
            launch();
        }
        g_launch_timing = LaunchTiming{};
        HIPCHECK(hipEventSynchronize(ev.back()));
        double sum = 0.0;
        for (int r = 0; r < reps; ++r) {
            float t = 0.f;
            HIPCHECK(hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
            sum += t;
        }
        *avg_ms = sum / reps;
        return FEM_OK;
    });
}

int fem_get_phase_times(fem_ctx* c, fem_phase_times* out) {
    return guarded(c, [&] {
        if (!out) throw InvalidArg("null output");
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->timer.resolve();
        for (int i = 0; i < FEM_NUM_PHASES; ++i) {
            out->ms[i] = c->timer.ms[i];
            out->launches[i] = c->timer.launches[i];
        }
        return FEM_OK;
    });
}

int64_t fem_num_nonzeros(fem_ctx* c) { return (c && c->mesh_ready) ? 9 * c->num_blocks : 0; }

int fem_hex_distinct_matrices(fem_ctx* c) {
    return (c && c->mesh_ready && c->kind == FEM_ELEMENT_HEX) ? c->hex.num_distinct : 0;
}

int fem_set_instances(fem_ctx* c, int32_t num_instances, const int64_t* node_offsets, const double* dt) {
    return guarded(c, [&] {
        if (!c->mesh_ready) return fail(c, FEM_ERR_ORDER, "upload the mesh before declaring instances");
        if (c->kind == FEM_ELEMENT_HEX) throw InvalidArg("batched instances cover tetrahedral meshes");
        if (c->opt.linear_solver == FEM_LINEAR_SOLVER_LU)
            throw InvalidArg("batched instances solve with CG (the direct solver covers one representation)");
        if (num_instances < 1 || !node_offsets) throw InvalidArg("need at least one instance and node offsets");
        const int M = num_instances;
        if (node_offsets[0] != 0 || node_offsets[M] != c->num_nodes)
            throw InvalidArg("node offsets must start at 0 and end at the node count");
        for (int i = 0; i < M; ++i)
            if (node_offsets[i + 1] <= node_offsets[i]) throw InvalidArg("every instance needs at least one node");
        if (dt)
            for (int i = 0; i < M; ++i)
                if (!(dt[i] > 0.0)) throw InvalidArg("instance time steps must be positive");
        HIPCHECK(hipSetDevice(c->device));
        c->h_node_off.assign(node_offsets, node_offsets + M + 1);
        std::vector<int> ni(c->num_nodes), ei(c->num_elements);
        for (int i = 0; i < M; ++i)
            for (int64_t a = node_offsets[i]; a < node_offsets[i + 1]; ++a) ni[a] = i;
        std::vector<int4> conn(c->num_elements);
        HIPCHECK(hipMemcpy(conn.data(), c->conn4.p, conn.size() * sizeof(int4), hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < c->num_elements; ++e) {
            const int i = ni[conn[e].x];
            if (ni[conn[e].y] != i || ni[conn[e].z] != i || ni[conn[e].w] != i)
                throw InvalidArg("element " + std::to_string(e) + " joins two instances");
            ei[e] = i;
        }
        c->num_instances = M;
        c->linear_matrix_ready = false;
        c->elem_inst.alloc(c->num_elements);
        c->elem_inst.upload(ei.data(), c->num_elements, c->stream);
        c->node_inst.alloc(c->num_nodes);
        c->node_inst.upload(ni.data(), c->num_nodes, c->stream);
        c->b_inst_dt.alloc(M);
        if (dt) {
            c->h_inst_dt.assign(dt, dt + M);
            c->b_inst_dt.upload(dt, M, c->stream);
        } else {
            c->h_inst_dt.clear();
            c->b_dt_for = -1.0;
        }
        c->b_inst_maxit.alloc(M);
        c->b_maxit_for = -2.0;
        c->b_sc.alloc(M);
        c->b_outer.alloc(M);
        c->b_kA.alloc(M);
        c->b_kV.alloc(M);
        c->h_inst_active.assign(M, 1);
        c->h_inst_unconv.assign(M, 0);
        c->h_inst_fp64.assign(M, 0);
        c->b_skip32.alloc(M);
        c->b_skip32.upload(c->h_inst_fp64.data(), M, c->stream);
        c->b_skip64.alloc(M);
        c->last_precision = -1;  // the CG vectors are sized for the FP64 pass of a batch at the next step
        c->b_active.alloc(M);
        c->b_active.upload(c->h_inst_active.data(), M, c->stream);
        c->b_invalid.alloc(M);
        HIPCHECK(hipMemsetAsync(c->b_invalid.p, 0, M * sizeof(int), c->stream));
        if (c->h_bsc) HIPCHECK(hipHostFree(c->h_bsc));
        c->h_bsc = nullptr;
        HIPCHECK(hipHostMalloc((void**)&c->h_bsc, M * sizeof(CgScalars), hipHostMallocDefault));
        std::vector<CgScalars> zero(M);
        std::memset(zero.data(), 0, M * sizeof(CgScalars));
        c->b_sc.upload(zero.data(), M, c->stream);
        c->b_outer.upload(zero.data(), M, c->stream);
        c->h_inst_iters.assign(M, 0);
        c->h_inst_err.assign(M, 0.0);
        c->h_inst_conv.assign(M, 0);
        for (int i = 0; i < 17; ++i) c->b_predicted[i] = 8;
        build_batch_tables(c, c->b_fixed_rows > 0 ? c->b_fixed_rows : 256, c->b_fixed_lanes > 0 ? c->b_fixed_lanes : 8);
        c->b_tuned = false;
        HIPCHECK(hipStreamSynchronize(c->stream));
        return FEM_OK;
    });
}

int fem_query_instances(fem_ctx* c, fem_instance_stats* out) {
    return guarded(c, [&] {
        if (c->num_instances == 0) return fail(c, FEM_ERR_ORDER, "no instances declared");
        if (!out) throw InvalidArg("null output");
        for (int i = 0; i < c->num_instances; ++i) {
            out[i].last_iterations = c->h_inst_iters[i];
            out[i].last_error = c->h_inst_err[i];
            out[i].unconverged_steps = c->h_inst_unconv[i];
            out[i].solved_in_fp64 = c->h_inst_fp64[i];
            out[i].reserved = 0;
            out[i].converged = c->h_inst_conv[i];
            out[i].active = c->h_inst_active[i];
        }
        return FEM_OK;
    });
}

int fem_set_batch_tuning(fem_ctx* c, int32_t chunk_rows, int32_t lanes_per_row, int32_t use_graphs) {
    return guarded(c, [&] {
        if (chunk_rows < 0 || (chunk_rows > 0 && lanes_per_row != 4 && lanes_per_row != 8 && lanes_per_row != 16))
            throw InvalidArg("chunk_rows >= 0 and lanes_per_row in {4, 8, 16} (0, 0: autotune)");
        c->b_fixed_rows = chunk_rows;
        c->b_fixed_lanes = chunk_rows > 0 ? lanes_per_row : 0;
        c->b_tuned = false;
        c->b_use_graphs = use_graphs != 0;
        reset_batch_graphs(c);
        return FEM_OK;
    });
}

int fem_get_batch_tuning(fem_ctx* c, int32_t* chunk_rows, int32_t* lanes_per_row, int32_t* num_chunks,
                         double* ms_per_iteration) {
    return guarded(c, [&] {
        if (chunk_rows) *chunk_rows = c->b_chunk_rows;
        if (lanes_per_row) *lanes_per_row = c->b_lanes;
        if (num_chunks) *num_chunks = c->b_num_chunks;
        if (ms_per_iteration) *ms_per_iteration = c->b_tune_ms;
        return FEM_OK;
    });
}

int fem_set_young_modulus(fem_ctx* c, const double* young) {
    return guarded(c, [&] {
        if (!c->mesh_ready || !young) throw InvalidArg("mesh not uploaded or null input");
        HIPCHECK(hipSetDevice(c->device));
        std::vector<double> mat(4 * c->num_elements);
        for (int64_t e = 0; e < c->num_elements; ++e) {
            const double E = young[e], nu = c->h_nu[e];
            check_material(E, nu, c->h_rho[e], e);
            mat[4 * e] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
            mat[4 * e + 1] = E / (2.0 * (1 + nu));
            mat[4 * e + 2] = c->h_rho[e];
            mat[4 * e + 3] = E;
        }
        c->mat.upload(reinterpret_cast<const double4*>(mat.data()), c->num_elements, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->linear_matrix_ready = false;  // operators depend on E
        c->hex.a32_lumped = -1;
        return FEM_OK;
    });
}

namespace {

int adjoint_checks(fem_ctx* c) {
    if (!c->mesh_ready) return fail(c, FEM_ERR_ORDER, "mesh not uploaded");
    if (c->kind == FEM_ELEMENT_HEX) throw InvalidArg("the adjoint covers tetrahedral meshes");
    if (c->num_instances > 0) throw InvalidArg("the adjoint covers single (non-batched) meshes");
    if (c->opt.rayleigh_mass != 0.0 || c->opt.rayleigh_stiffness != 0.0 || c->opt.mass_lumping ||
        scheme_mode(c->opt.scheme) != 0)
        throw InvalidArg("the adjoint reverses undamped implicit Euler steps with the consistent mass");
    if (c->kind == FEM_ELEMENT_TET_COROTATIONAL &&
        (c->opt.scheme != FEM_SCHEME_EULER_IMPLICIT || c->opt.newton_max_iterations != 1))
        throw InvalidArg("the corotational adjoint reverses one-Newton-iteration implicit Euler steps");
    HIPCHECK(hipSetDevice(c->device));
    ensure_bc(c);
    alloc_precision_buffers(c);
    return FEM_OK;
}

// One adjoint step (the transpose of the implicit-Euler step from (x_n, v_n)): the operator at the step's
// start, S mu = lam_v + dt lam_x, then the element tangents and the gather of the new adjoints.
void adjoint_step_impl(fem_ctx* c, const double* x_n, const double* v_n, const double* v_next, const double* lam_x,
                       const double* lam_v, double* lam_x_out, double* lam_v_out, double* grad_young) {
    const size_t n = (size_t)c->ndof;
    hipStream_t s = c->stream;
    // operator S at the step's starting state (element kernel + assembly; the rhs is not needed)
    HIPCHECK(hipMemcpyAsync(c->xs.p, x_n, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(c->vs.p, v_n, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    {
        // element matrices evaluated at (x_n, v_n): the kernel's xn/vn are the evaluation point
        // (the reference's newState), xs/vs the step's start state; both are x_n, v_n here
        TetElementArgs ea;
        ea.num_elements = (int)c->num_elements;
        ea.x0 = c->x0.p;
        ea.xn = c->xs.p;
        ea.vn = c->vs.p;
        ea.xs = c->xs.p;
        ea.vs = c->vs.p;
        ea.dt = c->opt.dt;
        ea.write_matrix = 1;
        ea.general = 0;
        set_staging(c, ea);
        launch_tet_elements(c->kind, ea, s);
    }
    assemble_tet(c);
    // S mu = lv + dt lx (transpose solve on the symmetric operator)
    launch_adjoint_rhs((int)n, lam_x, lam_v, c->opt.dt, c->b.p, s);
    c->solving_adjoint = true;  // its own iteration predictor
    try {
        linear_solve(c);
    } catch (...) {
        c->solving_adjoint = false;
        throw;
    }
    c->solving_adjoint = false;
    if (!c->adj_gx.p) {
        c->adj_gx.alloc(12 * (size_t)c->num_elements);
        c->adj_gv.alloc(12 * (size_t)c->num_elements);
    }
    AdjointArgs a;
    a.num_elements = (int)c->num_elements;
    a.conn = c->conn4.p;
    a.x0 = c->x0.p;
    a.x = c->xs.p;
    a.v = c->vs.p;
    a.v_next = v_next;
    a.mu = c->dv.p;
    a.bcmask = c->bcmask.p;
    a.mat = c->mat.p;
    a.dt = c->opt.dt;
    a.gx_rec = c->adj_gx.p;
    a.gv_rec = c->adj_gv.p;
    a.grad_E = grad_young;
    launch_adjoint_elements(c->kind, a, s);
    launch_adjoint_gather((int)c->num_nodes, c->node_ptr.p, c->node_list.p, c->adj_gx.p, c->adj_gv.p, lam_x,
                          c->b.p, lam_x_out, lam_v_out, s);
    HIPCHECK(hipGetLastError());
}

// nsteps adjoint steps backwards through a history (rows of ndof: states 0 .. nsteps), lam_x, lam_v updated
// in place: step j = nsteps - 1 .. 0 reverses the step from state j to state j + 1.
// The call's first transpose solve starts from zero; each later one from the solution of the step after it
// (still in dv: nothing between two adjoint solves writes it), which lies close to its own. The result of a call
// depends only on its arguments.
void adjoint_loop(fem_ctx* c, int32_t nsteps, const double* x_hist, const double* v_hist, double* lam_x,
                  double* lam_v, double* grad_young) {
    const size_t n = (size_t)c->ndof;
    for (int j = nsteps - 1; j >= 0; --j) {
        c->adj_warm = j < nsteps - 1;
        try {
            adjoint_step_impl(c, x_hist + j * n, v_hist + j * n, v_hist + (j + 1) * n, lam_x, lam_v, lam_x, lam_v,
                              grad_young);
        } catch (...) {
            c->adj_warm = false;
            throw;
        }
    }
    c->adj_warm = false;
}

}  // namespace

int fem_adjoint_step(fem_ctx* c, const double* x_n, const double* v_n, const double* v_next, const double* lam_x,
                     const double* lam_v, double* lam_x_out, double* lam_v_out, double* grad_young) {
    return guarded(c, [&] {
        if (!x_n || !v_n || !v_next || !lam_x || !lam_v || !lam_x_out || !lam_v_out || !grad_young)
            throw InvalidArg("null device pointer");
        const int rc = adjoint_checks(c);
        if (rc != FEM_OK) return rc;
        adjoint_step_impl(c, x_n, v_n, v_next, lam_x, lam_v, lam_x_out, lam_v_out, grad_young);
        return FEM_OK;
    });
}

int fem_adjoint_steps(fem_ctx* c, int32_t nsteps, const double* x_hist, const double* v_hist, double* lam_x,
                      double* lam_v, double* grad_young) {
    return guarded(c, [&] {
        if (!x_hist || !v_hist || !lam_x || !lam_v || !grad_young) throw InvalidArg("null device pointer");
        if (nsteps < 0) throw InvalidArg("negative step count");
        const int rc = adjoint_checks(c);
        if (rc != FEM_OK) return rc;
        adjoint_loop(c, nsteps, x_hist, v_hist, lam_x, lam_v, grad_young);
        return FEM_OK;
    });
}

int fem_get_node_masses(fem_ctx* c, double* mass) {
    return guarded(c, [&] {
        if (!c->mesh_ready || !mass) throw InvalidArg("mesh not uploaded or null output");
        std::copy(c->h_mass_per_node.begin(), c->h_mass_per_node.end(), mass);
        return FEM_OK;
    });
}

int fem_get_pattern(fem_ctx* c, int32_t* rowptr, int32_t* colidx) {
    return guarded(c, [&] {
        if (!c->mesh_ready) throw InvalidArg("mesh not uploaded");
        const auto& rp = c->h_rowptr_b;
        const auto& ci = c->h_colidx_b;
        int64_t pos = 0;
        for (int64_t I = 0; I < c->num_nodes; ++I) {
            int64_t len = rp[I + 1] - rp[I];
            for (int r = 0; r < 3; ++r) {
                if (rowptr) rowptr[3 * I + r] = (int32_t)pos;
                if (colidx)
                    for (int64_t k = rp[I]; k < rp[I + 1]; ++k)
                        for (int cc = 0; cc < 3; ++cc) colidx[pos + 3 * (k - rp[I]) + cc] = 3 * ci[k] + cc;
                pos += 3 * len;
            }
        }
        if (rowptr) rowptr[c->ndof] = (int32_t)pos;
        return FEM_OK;
    });
}

int fem_get_system_matrix(fem_ctx* c, double* values) {
    return guarded(c, [&] {
        if (!c->mesh_ready || !values) throw InvalidArg("mesh not uploaded or null output");
        if (c->kind == FEM_ELEMENT_HEX) throw InvalidArg("hexahedral meshes use the element-by-element operator");
        HIPCHECK(hipSetDevice(c->device));
        std::vector<double> S(9 * c->num_blocks);
        HIPCHECK(hipMemcpy(S.data(), c->S64.p, S.size() * sizeof(double), hipMemcpyDeviceToHost));
        const auto& rp = c->h_rowptr_b;
        int64_t pos = 0;
        for (int64_t I = 0; I < c->num_nodes; ++I)
            for (int r = 0; r < 3; ++r)
                for (int64_t k = rp[I]; k < rp[I + 1]; ++k)
                    for (int cc = 0; cc < 3; ++cc) values[pos++] = S[9 * k + 3 * r + cc];
        return FEM_OK;
    });
}

int fem_get_rhs(fem_ctx* c, double* rhs) {
    return guarded(c, [&] {
        if (!c->mesh_ready || !rhs) throw InvalidArg("mesh not uploaded or null output");
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipMemcpy(rhs, c->b.p, (size_t)c->ndof * sizeof(double), hipMemcpyDeviceToHost));
        return FEM_OK;
    });
}

int fem_evaluate_elements(fem_ctx* c, const double* x, int64_t first, int64_t count, double* R, double* K,
                          double* f) {
    return guarded(c, [&] {
        if (!c->mesh_ready || !x) throw InvalidArg("mesh not uploaded or null positions");
        if (first < 0 || count < 0 || first + count > c->num_elements) throw InvalidArg("element range out of bounds");
        HIPCHECK(hipSetDevice(c->device));
        const int nn3 = 3 * c->nn;
        DevBuf<double> dx, dR, dK, df;
        dx.alloc((size_t)c->ndof);
        dx.upload(x, (size_t)c->ndof, c->stream);
        dR.alloc((size_t)count * 9);
        dK.alloc((size_t)count * nn3 * nn3);
        df.alloc((size_t)count * nn3);
        if (c->kind == FEM_ELEMENT_HEX)
            hex_eval_elements(c->hex, (int)first, (int)count, c->x0.p, dx.p, c->mat.p, dR.p, dK.p, df.p, c->stream);
        else
            launch_eval_elements(c->kind, (int)first, (int)count, c->conn4.p, c->x0.p, dx.p, c->mat.p, dR.p, dK.p,
                                 df.p, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        HIPCHECK(hipGetLastError());
        if (R) HIPCHECK(hipMemcpy(R, dR.p, (size_t)count * 9 * sizeof(double), hipMemcpyDeviceToHost));
        if (K) HIPCHECK(hipMemcpy(K, dK.p, (size_t)count * nn3 * nn3 * sizeof(double), hipMemcpyDeviceToHost));
        if (f) HIPCHECK(hipMemcpy(f, df.p, (size_t)count * nn3 * sizeof(double), hipMemcpyDeviceToHost));
        return FEM_OK;
    });
}

}  // extern "C"
