// Hexahedral (Fem3DElementCube) path: element-by-element operator.
//
// Reference semantics (paths relative to the reference root):
//   element K, M (2x2x2 Gauss)    Physics/Fem3DElementCube.cpp:65-366, Math/GaussLegendreQuadrature.cpp:30-34
//   force f = -K (x - x0)         Physics/Fem3DElementCube.cpp:129-141
//   element-by-element products   Physics/FemElement.cpp:168-210 (addMatVec)
//   implicit step                 Math/OdeSolverEulerImplicit.cpp:55-161
// K and M of a trilinear hex are constant, so the system matrix A = M/dt + dt K is constant too.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../../include/fem_gpu.h"
#include "fem_kernels.hpp"

namespace femgpu {

struct HexData {
    int64_t num_nodes = 0;
    int64_t num_elements = 0;
    int num_colors = 0;
    std::vector<int> color_ptr;   // elements of color c: order[color_ptr[c] .. color_ptr[c + 1])
    int* conn = nullptr;          // [E][8] node ids
    int* order = nullptr;         // elements sorted by color
    std::vector<int> pad_ptr;     // colour c: order_pad[pad_ptr[c] .. pad_ptr[c + 1]), multiples of 64
    int* order_pad = nullptr;     // colour order padded to 64-element tiles (-1 = padding)
    float* A32 = nullptr;         // packed upper triangle of M/dt + dt K (FP32): 28 upper node blocks of 9 and
                                  // 8 symmetric diagonal blocks of 6, 300 floats per padded position,
                                  // interleaved by 64-position tiles in float4 chunks
    double* diag64 = nullptr;     // diagonal of A (per DOF)
    float* diag32 = nullptr;
    // Distinct cached matrices (a structured mesh repeats a few element matrices): when every padded
    // position's 300 floats equal one of at most kHexMaxDistinct representatives bit for bit, the FP32
    // operator reads table[u] (75 float4 chunks each, staged in LDS) through table_idx instead of A32.
    float* table = nullptr;             // [num_distinct][300]
    int* table_idx = nullptr;           // [padded positions] table entry of each position
    int num_distinct = 0;               // 0: A32 per position
    std::vector<int> h_order_pad;       // host copy of order_pad (padding positions are -1)
    double* mdiag_cons = nullptr;       // diag(M) per DOF, consistent and lumped (row sums); built on first use
    double* mdiag_lump = nullptr;
    double* diagx64 = nullptr;          // diag64 plus the localized external diagonal of a step that has one
    float* diagx32 = nullptr;
    double a32_cm = 0.0, a32_ck = 0.0;  // coefficients of M and K the cached A32 / diagonal were built for
    int a32_lumped = -1;                // lumped mass flag of the cache (-1: not built)
    bool ready = false;
};

// Host set-up: element validity (positive volume, invertible J, Physics/Fem3DElementCube.cpp:169-211, 321-366),
// lumped node masses (Physics/FemRepresentation.cpp:102-117), the node-pair block pattern and the
// element colouring. mat: per element (lambda, mu, rho, E).
void hex_setup(HexData& h, int64_t num_nodes, const double* x0, int64_t num_elements, const int32_t* conn,
               const double* mat, std::vector<double>& mass_per_node, std::vector<int32_t>& rowptr_b,
               std::vector<int32_t>& colidx_b, hipStream_t s);
void hex_release(HexData& h);

// One implicit step on the hexahedral mesh. Work buffers come from the caller.
struct HexStepArgs {
    const fem_options* opt;
    const double* x0;
    double* x;
    double* v;
    const double* fg;       // gravity m_i g
    const double* fext;     // may be null
    const double* fstep = nullptr;  // localized force of this step (may be null)
    const uint8_t* bcmask;
    const double4* mat;
    double* b;              // rhs
    double* dv;             // solution
    double* r64;
    void* vec[7];           // x, r, z, p, q, dinv, spare (TV = float or double)
    double* partial;
    CgScalars* sc;          // [2]
    CgScalars* h_sc;        // pinned host copies [2]
    int* invalid;
    int max_iter;
    int* predicted;         // [17]: FP64 solve, then per refinement sweep
    // Localized external stiffness and damping of the step (Physics/Fem3DRepresentation.cpp:133-198): operator
    // rows (row DOF, then column DOFs and values in ext_ptr[r] .. ext_ptr[r + 1]) under the lower-triangle view,
    // added to every application of the operator; their diagonal entries join the Jacobi diagonal. Boundary
    // rows and columns carry none.
    int ext_rows = 0;
    const int* ext_row = nullptr;
    const int* ext_ptr = nullptr;
    const int* ext_col = nullptr;
    const double* ext_val = nullptr;
    const float* ext_val32 = nullptr;
    // Right-hand-side rows of implicit Euler, -dt K_ext v with the full external K (fval zero: the force itself
    // is in fstep), in launch_ext_rhs's layout.
    int rhs_rows = 0;
    const int* rhs_row = nullptr;
    const int* rhs_ptr = nullptr;
    const int* rhs_col = nullptr;
    const double* rhs_k = nullptr;
    const double* rhs_f = nullptr;
    const double* rhs_dt = nullptr;
};
void hex_step(HexData& h, const HexStepArgs& a, fem_stats& stats, hipStream_t s);
// The linear solve of hex_step alone: dv = A^-1 b with the operator cached by the last step (b's boundary
// rows already zero).
void hex_solve(HexData& h, const HexStepArgs& a, fem_stats& stats, hipStream_t s);

// The operator the solves apply: A = cm M + ck K (lumped: M's row sums on the diagonal); the FP32 cache and the
// diagonals are rebuilt only when the coefficients change.
void hex_set_operator(HexData& h, const HexStepArgs& a, double cm, double ck, int lumped, hipStream_t s);
// a.b = (fg + fext + fstep - K ((x - x0) + kv v) - am M v) / div with boundary rows zero: the force f(x, v) of the
// schemes (kv = alpha_K, am = alpha_M: the Rayleigh force -D v).
// With damp_d (diag(M) per DOF): the Rayleigh mass term as -am diag(M) v per DOF instead of -am M v.
void hex_rhs(HexData& h, const HexStepArgs& a, const double* x, const double* v, double kv, double am, int lumped,
             double div, hipStream_t s, const double* damp_d = nullptr);
// diag(M) per DOF (lumped: M's row sums), built on the first call.
const double* hex_mass_diagonal(HexData& h, const double* x0, const double4* mat, bool lumped, hipStream_t s);

// Average milliseconds of one application of the element-by-element operator over every colour (q += A p on
// the CG scratch vectors p = vec[3], q = vec[4]): the FP32 cached operator (table or per-element cache, as the
// solves use it) or the FP64 matrix-free one; reps applications, each between its own pair of events.
double hex_time_operator(HexData& h, const HexStepArgs& a, bool fp32, int reps, hipStream_t s);

// FP64 element outputs for tests: K (24x24 row-major) and f = -K (x - x0).
void hex_eval_elements(HexData& h, int first, int count, const double* x0, const double* x, const double4* mat,
                       double* R, double* K, double* f, hipStream_t s);

}  // namespace femgpu
