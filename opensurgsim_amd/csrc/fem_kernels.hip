// gfx950 kernels of the Fem3D implicit step: tetrahedron element kernels, deterministic gather
// assembly into a 3x3-block operator, right-hand side, Jacobi-preconditioned CG and the
// implicit-Euler state update.
//
// Reference semantics (paths relative to the reference root):
//   element kernels  Physics/Fem3DElementCorotationalTetrahedron.cpp:77-288, Physics/Fem3DElementTetrahedron.cpp:77-216
//   assembly         Physics/FemRepresentation.cpp:479-545 (values), Math/SparseMatrix.h:178-242
//   system matrix    Math/OdeSolverEulerImplicit.cpp:148-160 (A = M/dt + D + dt K, rhs) with
//                    Math/OdeState.cpp:159-218 (boundary conditions)
//   CG               Eigen::ConjugateGradient<SparseMatrix<double>> (Lower, DiagonalPreconditioner)
//                    behind Math/LinearSparseSolveAndInverse.cpp:119-129
//   update           Math/OdeSolverEulerImplicit.cpp:98-99
#include "fem_device.hpp"
#include "fem_kernels.hpp"
#include "fem_reduce.hpp"

#include <algorithm>
#include <cfloat>
#include <stdexcept>

namespace femgpu {

thread_local LaunchTiming g_launch_timing;


// ------------------------------------------------------------------------------------------
// Tetrahedron element kernel (FP64). Four lanes per element, lane a for node a: every lane forms the
// element's rotation (the element-wide part is cheap next to the stiffness blocks), then its own row:
// f_a, the blocks K_ab (b = 0..3), the right-hand-side row f_a + sum_b K_ab (xn - xs - vn dt)_b
// - sum_b M_ab (vn - vs)_b / dt, and its lower-triangle blocks of A_e = M_e / dt + dt K_e (the diagonal
// block a, 6 values, and the pair blocks (a, b) whose node id of a is the larger, 9 values). Holding one
// row instead of the whole element keeps the registers to a few waves per SIMD.
// A workgroup is a group of 64 consecutive elements (kElemGroup). Its lanes put their contributions in LDS;
// then the group sums, in element order, the contributions its elements make to each lower block and to each
// node's rhs row (a cube's five tets share most of their node pairs, neighbouring cubes some more), and stores
// one record per (group, block) and per (group, node) where the consumer reads it: in the staging list ordered
// by block (by node), records in group order, which is element order. The assembly and the right-hand side
// then stream those lists: on the 40^3 block 2.5x fewer bytes of block records (85 against 215 MB) and 4.4x
// fewer rhs records than one record per element.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ constexpr int pair_slot(int a, int b) { return a * (7 - a) / 2 + (b - a - 1); }

// XCD-aware logical workgroup index: workgroups are dispatched round-robin over the 8 XCDs, so with a grid
// that is a multiple of 8, L = (b % 8) * (grid / 8) + b / 8 gives every XCD one contiguous slab of the index
// space (neighbouring elements, nodes or blocks share their gathers in that XCD's L2).
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x;
    return (nb % kXcds == 0) ? (blockIdx.x % kXcds) * (nb / kXcds) + blockIdx.x / kXcds : blockIdx.x;
}
inline dim3 xcd_grid(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    b = (b + kXcds - 1) / kXcds * kXcds;
    return dim3((unsigned)b);
}

// v[a] for a lane-dependent a without dynamic indexing of a register array (which would go to scratch).
template <int N>
__device__ __forceinline__ void pick(const double (&v)[4][N], int a, double* out) {
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = a == 0 ? v[0][i] : a == 1 ? v[1][i] : a == 2 ? v[2][i] : v[3][i];
}

// out = v of lane b of this lane's quad (DPP quad_perm [b, b, b, b]); the whole quad must be active.
__device__ __forceinline__ void quad_bcast3(int b, const double* v, double* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        switch (b) {
            case 0: out[i] = dpp_mov<0x00>(v[i]); break;
            case 1: out[i] = dpp_mov<0x55>(v[i]); break;
            case 2: out[i] = dpp_mov<0xAA>(v[i]); break;
            default: out[i] = dpp_mov<0xFF>(v[i]); break;
        }
    }
}

// Lane a of element e: its row of the element's contributions, into lm (the element's 10 slots of 9 values in
// LDS: pairs (0,1) .. (2,3) oriented with the larger node id as row, each written by that node's lane, then
// the diagonal blocks' lower triangles) and lr (its 4 rhs rows of 3).
template <bool COROT, bool GENERAL>
__device__ __forceinline__ void tet_element_row(const TetElementArgs& A, int e, int a, double* lm, double* lr) {
    const int4 id4 = A.conn[e];
    const int ids[4] = {id4.x, id4.y, id4.z, id4.w};
    const int ida = a == 0 ? id4.x : a == 1 ? id4.y : a == 2 ? id4.z : id4.w;
    const double dt = A.elem_inst ? A.inst_dt[A.elem_inst[e]] : A.dt;
    const double4 m = A.mat[e];
    const double lam = m.x, mu = m.y, rho = m.z;
    double V, ga[3], ra[3];
    double R[9], W[9], ha[3];
    bool correction = false;
    {
        double x0[12], xn[12];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const size_t i = 3 * (size_t)ids[b] + c;
                x0[3 * b + c] = A.x0[i];
                xn[3 * b + c] = A.xn[i];
            }
        double g[4][3];
        V = tet_rest_gradients(x0, g);
        pick<3>(g, a, ga);
        if (COROT) {
            CoroCommon<double> cc;
            coro_common<double>(x0, xn, g, lam, mu, cc);
            coro_force<double>(cc, ga, V, ra);
            if (cc.flags && a == 0) {
                if (cc.flags & ELEM_FLAG_DET_R) atomicAdd(&A.flag_counts[0], 1);
                if (cc.flags & ELEM_FLAG_G_SINGULAR) atomicAdd(&A.flag_counts[1], 1);
                if (cc.flags & ELEM_FLAG_POLAR_EIG) atomicAdd(&A.flag_counts[2], 1);
            }
            correction = cc.correction;
#pragma unroll
            for (int i = 0; i < 9; ++i) R[i] = cc.R[i];
            mv3(cc.R, ga, ha);
            coro_weight<double>(cc, ga, V, lam, mu, W);
        } else {
            // f_a = -sum_b K0_ab (x_b - x0_b)  (Physics/Fem3DElementTetrahedron.cpp:77-90)
            double tt[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double Kb[9];
                linear_row_block(ga, g[b], V, lam, mu, Kb);
                double u[3] = {xn[3 * b] - x0[3 * b], xn[3 * b + 1] - x0[3 * b + 1], xn[3 * b + 2] - x0[3 * b + 2]};
#pragma unroll
                for (int i = 0; i < 3; ++i) tt[i] += Kb[3 * i] * u[0] + Kb[3 * i + 1] * u[1] + Kb[3 * i + 2] * u[2];
            }
            ra[0] = -tt[0];
            ra[1] = -tt[1];
            ra[2] = -tt[2];
        }
    }
    // Mass: M_e = R M0 R^T = rho V / 20 (2 on the diagonal blocks, 1 off) x I3; lumped: the row sum
    // 5 rho V / 20 on the diagonal blocks (Physics/FemRepresentation.cpp:499-508).
    const double mc = rho * V / 20.0;
    const double idt = 1.0 / dt;
    // operator coefficients of M and K, and the coefficient of K vn in the rhs, by scheme
    // (D = alpha_M M + alpha_K K, Physics/FemRepresentation.cpp:510-518, 565-629)
    const double cm = A.mode == 0 ? idt + A.ray_mass : (A.mode == 1 ? idt : 0.0);
    const double ck = A.mode == 0 ? dt + A.ray_stiff : (A.mode == 1 ? 0.0 : 1.0);
    const double kv = A.mode == 0 ? dt + A.ray_stiff : A.ray_stiff;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double gb[3];
        quad_bcast3(b, ga, gb);  // g_b and R g_b from lane b of the quad
        double K[9];
        if (COROT) {
            double hb[3];
            quad_bcast3(b, ha, hb);
            coro_row_block(ha, hb, ga, gb, W, R, correction, V, lam, mu, K);
        } else {
            linear_row_block(ga, gb, V, lam, mu, K);
        }
        // rhs = f + K (xn - xs - vn dt) - M (vn - vs) / dt   (Math/OdeSolverEulerImplicit.cpp:157-158)
        double d1[3], d2[3], vb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t i = 3 * (size_t)ids[b] + c;
            const double vv = A.vn[i];
            vb[c] = vv;
            if (GENERAL) {
                d1[c] = (A.xn[i] - A.xs[i]) - vv * kv;
                d2[c] = vv - A.vs[i];
            } else {
                d1[c] = -(vv * kv);
                d2[c] = 0.0;
            }
        }
        const double mab = A.lumped ? (a == b ? 5.0 * mc : 0.0) : ((a == b) ? 2.0 * mc : mc);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ra[i] += K[3 * i] * d1[0] + K[3 * i + 1] * d1[1] + K[3 * i + 2] * d1[2];
            if (GENERAL) ra[i] -= (mab * d2[i]) / dt;
            if (A.ray_mass != 0.0) ra[i] -= A.ray_mass * (mab * vb[i]);
        }
        if (!A.write_matrix) continue;
        // A_ab = M_ab cm + K_ab ck   (implicit: M / dt + D + dt K, Math/OdeSolverEulerImplicit.cpp:148; the
        // elements' own damping is zero for 3D elements, Physics/FemElement.cpp:131-145)
        const double mdt = mab * cm;
        if (a == b) {
            double* o = lm + 9 * (6 + a);
            o[0] = mdt + K[0] * ck;
            o[1] = K[3] * ck;
            o[2] = mdt + K[4] * ck;
            o[3] = K[6] * ck;
            o[4] = K[7] * ck;
            o[5] = mdt + K[8] * ck;
        } else if (ida > ids[b]) {
            const int ps = (a < b) ? pair_slot(a, b) : pair_slot(b, a);
            double* o = lm + 9 * ps;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) o[3 * i + j] = ((i == j) ? mdt : 0.0) + K[3 * i + j] * ck;
        }
    }
    lr[3 * a] = ra[0];
    lr[3 * a + 1] = ra[1];
    lr[3 * a + 2] = ra[2];
}

template <bool COROT, bool GENERAL>
__global__ __launch_bounds__(256) void tet_element_kernel(TetElementArgs A) {
    static_assert(4 * kElemGroup == 256, "one workgroup is one element group");
    constexpr int kMaxSrc = 10 * kElemGroup, kMaxNodeSrc = 4 * kElemGroup;  // contributions of a group
    // contributions of the group's elements: per element its 10 slots of 9 values (the diagonal blocks use the
    // first 6) and its 4 rhs rows
    __shared__ double cmat[kElemGroup * 10 * 9];
    __shared__ double crhs[kElemGroup * 4 * 3];
    __shared__ uint16_t ssrc[kMaxSrc];
    __shared__ uint16_t snsrc[kMaxNodeSrc];
    const int g = xcd_block();
    const int tid = (int)threadIdx.x;
    const int64_t t = (int64_t)g * blockDim.x + tid;
    const int e = (int)(t >> 2), a = (int)(t & 3), le = tid >> 2;
    // The group's record tables are read before the element work, so that their latency hides behind it:
    // thread tid takes block record tid and node record tid (the first of its records) and sources tid,
    // tid + 256, tid + 512 of the group's source lists.
    const bool grp = g < A.num_groups;  // false in the padding workgroups of the XCD-rounded grid
    int q0 = 0, q1 = 0, n0 = 0, n1 = 0, sb = 0, nsb = 0;
    int qs = 0, qe = 0, qo = 0, ns = 0, ne = 0, no = 0;
    uint16_t sv[3] = {0, 0, 0}, nv = 0;
    if (grp) {
        n0 = A.grp_n[g];
        n1 = A.grp_n[g + 1];
        nsb = A.n_src[n0];
        if (n0 + tid < n1) {
            ns = A.n_src[n0 + tid];
            ne = A.n_src[n0 + tid + 1];
            no = A.n_out[n0 + tid];
        }
        if (nsb + tid < A.n_src[n1]) nv = A.nsrc16[nsb + tid];
        if (A.write_matrix) {
            q0 = A.grp_q[g];
            q1 = A.grp_q[g + 1];
            sb = A.q_src[q0];
            const int se = A.q_src[q1];
            if (q0 + tid < q1) {
                qs = A.q_src[q0 + tid];
                qe = A.q_src[q0 + tid + 1];
                qo = A.q_out[q0 + tid];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
                if (sb + tid + 256 * i < se) sv[i] = A.src16[sb + tid + 256 * i];
        }
    }
    if (e < A.num_elements) {  // the four lanes of an element take the branch together
        tet_element_row<COROT, GENERAL>(A, e, a, cmat + 90 * le, crhs + 12 * le);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (tid + 256 * i < kMaxSrc) ssrc[tid + 256 * i] = sv[i];
    snsrc[tid] = nv;
    __syncthreads();
    if (!grp) return;
    if (A.write_matrix) {
        // one lane per (group, block) record: the block's contributions summed in element order
        for (int q = q0 + tid; q < q1; q += 256) {
            int s0 = qs, s1 = qe, o = qo;
            if (q != q0 + tid) {
                s0 = A.q_src[q];
                s1 = A.q_src[q + 1];
                o = A.q_out[q];
            }
            const int nc = ssrc[s0 - sb] % 10 >= 6 ? 6 : 9;  // a diagonal block's record has 6 values
            double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int k = s0 - sb; k < s1 - sb; ++k) {
                const double* c = cmat + 9 * ssrc[k];
#pragma unroll
                for (int i = 0; i < 9; ++i)
                    if (i < nc) acc[i] += c[i];
            }
            double* out = A.rec + o;
#pragma unroll
            for (int i = 0; i < 9; ++i)
                if (i < nc) out[i] = acc[i];
        }
    }
    // one lane per (group, node) rhs record
    for (int q = n0 + tid; q < n1; q += 256) {
        int s0 = ns, s1 = ne, o = no;
        if (q != n0 + tid) {
            s0 = A.n_src[q];
            s1 = A.n_src[q + 1];
            o = A.n_out[q];
        }
        double acc[3] = {0.0, 0.0, 0.0};
        for (int k = s0 - nsb; k < s1 - nsb; ++k) {
            const double* c = crhs + 3 * snsrc[k];
            acc[0] += c[0];
            acc[1] += c[1];
            acc[2] += c[2];
        }
        A.rhs_rec[o] = acc[0];
        A.rhs_rec[o + 1] = acc[1];
        A.rhs_rec[o + 2] = acc[2];
    }
}

void launch_tet_elements(int kind, const TetElementArgs& a, hipStream_t s) {
    if (a.num_elements == 0) return;
    if (!a.conn || !a.mat || !a.rhs_rec || !a.grp_n || (a.write_matrix && (!a.rec || !a.grp_q)))
        throw std::invalid_argument("element kernel launched without its inputs or staging tables");
    const dim3 grid = xcd_grid(4 * (int64_t)a.num_elements, 256), block(256);
    if (kind == 1) {
        if (a.general) launch_k((tet_element_kernel<true, true>), grid, block, 0, s, a);
        else launch_k((tet_element_kernel<true, false>), grid, block, 0, s, a);
    } else {
        if (a.general) launch_k((tet_element_kernel<false, true>), grid, block, 0, s, a);
        else launch_k((tet_element_kernel<false, false>), grid, block, 0, s, a);
    }
}

// sum_{c = c0}^{c1 - 1} p[stride * c], added in order; loads go out four at a time so one lane keeps
// several in flight (the adds stay sequential, so the rounding is that of the plain loop).
template <int U = 4>
__device__ __forceinline__ double sum_strided(const double* __restrict__ p, int stride, int c0, int c1) {
    double acc = 0.0;
    for (int c = c0; c < c1; c += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = c + u < c1 ? p[stride * (int64_t)(c + u)] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (c + u < c1) acc += v[u];  // in contribution order
    }
    return acc;
}

// ------------------------------------------------------------------------------------------
// Deterministic assembly from the staged contributions: one lane per value of a lower block (9 lanes per
// off-diagonal block, 6 per diagonal block; the off-diagonal blocks come first). A lane adds its value
// over the block's contributions in element order, applies the boundary conditions (zero BC rows and
// columns, 1 on the BC diagonal, Math/OdeState.cpp:197-218) and writes it and its transposed mirror,
// giving the lower-triangle self-adjoint view that Eigen's CG applies
// (Math/LinearSparseSolveAndInverse.h:122). Contributions of off-diagonal block t are the kPairRec-double
// records ctr_ptr[t] .. ctr_ptr[t + 1] of stage; those of the diagonal blocks follow at
// kPairRec * nctr_off as kDiagRec-double records (lower triangle (0,0), (1,0), (1,1), (2,0), (2,1), (2,2)).
// lb_info[t] = (block, mirror block, boundary bits: DOF k of the row in bit k, of the column in bit 3 + k).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void assemble_kernel(int num_off, int num_diag, const int4* __restrict__ lb_info,
                                                       const int* __restrict__ ctr_ptr, int nctr_off,
                                                       const double* __restrict__ stage,
                                                       double* __restrict__ S64, float* __restrict__ S32) {
    // The diagonal blocks' lanes (about 4x the contributions of an off-diagonal block on a tet mesh) take the
    // first workgroups, so that their longer chains of loads overlap the off-diagonal lanes instead of trailing.
    const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const int64_t n_diag_lanes = 6 * (int64_t)num_diag;
    int lt, i, j;
    bool diag;
    double acc;
    if (t >= n_diag_lanes) {
        const int64_t u = t - n_diag_lanes;
        if (u >= 9 * (int64_t)num_off) return;
        lt = (int)(u / 9);
        const int v = (int)(u - 9 * (int64_t)lt);
        i = v / 3;
        j = v - 3 * i;
        diag = false;
        acc = sum_strided<8>(stage + v, kPairRec, ctr_ptr[lt], ctr_ptr[lt + 1]);
    } else {
        const int d = (int)(t / 6);
        const int v = (int)(t - 6 * (int64_t)d);
        lt = num_off + d;
        i = v == 0 ? 0 : (v < 3 ? 1 : 2);
        j = v == 0 ? 0 : (v < 3 ? v - 1 : v - 3);
        diag = true;
        acc = sum_strided<8>(stage + kPairRec * (int64_t)nctr_off + v, kDiagRec, ctr_ptr[lt] - nctr_off,
                             ctr_ptr[lt + 1] - nctr_off);
    }
    const int4 info = lb_info[lt];
    double val = acc;
    if (((info.z >> i) | (info.z >> (3 + j))) & 1) val = (diag && i == j) ? 1.0 : 0.0;
    S64[9 * (int64_t)info.x + 3 * i + j] = val;
    if (S32) S32[9 * (int64_t)info.x + 3 * i + j] = (float)val;
    if (!diag || i != j) {  // the mirror: block (J, I) value (j, i), or the upper entry of a diagonal block
        S64[9 * (int64_t)info.y + 3 * j + i] = val;
        if (S32) S32[9 * (int64_t)info.y + 3 * j + i] = (float)val;
    }
}

void launch_assemble(int num_off, int num_diag, const int4* lb_info, const int* ctr_ptr, int nctr_off,
                     const double* stage, double* S64, float* S32, hipStream_t s) {
    const int64_t lanes = 9 * (int64_t)num_off + 6 * (int64_t)num_diag;
    if (lanes == 0) return;
    launch_k(assemble_kernel, xcd_grid(lanes, 256), dim3(256), 0, s, num_off, num_diag, lb_info, ctr_ptr,
             nctr_off, stage, S64, S32);
}

// Right-hand side, one lane per DOF: the node's staged element rows in element order, gravity m_i g
// (Physics/FemRepresentation.cpp:641-662), external force; zero on BC DOFs (Math/OdeState.cpp:159-172).
// With damp_d: the Rayleigh mass damping force as the global form with a lumped mass takes it,
// -alpha_M diag(M) v (Physics/FemRepresentation.cpp:583-587), damp_d = diag(M) per DOF.
__global__ __launch_bounds__(256) void rhs_kernel(int num_nodes, const int* __restrict__ node_ptr,
                                                  const double* __restrict__ stage, const double* __restrict__ fg,
                                                  const double* __restrict__ fext, const uint8_t* __restrict__ bcmask,
                                                  double* __restrict__ b, double div, const double* __restrict__ damp_d,
                                                  const double* __restrict__ damp_v, double damp_alpha,
                                                  const int* __restrict__ node_inst,
                                                  const double* __restrict__ inst_div) {
    const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= 3 * (int64_t)num_nodes) return;
    const int I = (int)(i / 3), k = (int)(i - 3 * (int64_t)I);
    const double acc = sum_strided(stage + k, kRhsRec, node_ptr[I], node_ptr[I + 1]);
    double v = acc + fg[i];
    if (damp_d) v -= damp_alpha * (damp_d[i] * damp_v[i]);
    if (fext) v += fext[i];
    b[i] = bcmask[i] ? 0.0 : v / (inst_div ? inst_div[node_inst[I]] : div);
}

__global__ __launch_bounds__(256) void ext_operator_kernel(int count, const int* __restrict__ blk,
                                                           const double* __restrict__ add, double* __restrict__ S64,
                                                           float* __restrict__ S32) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * (int64_t)count) return;
    const size_t o = 9 * (size_t)blk[t / 9] + (size_t)(t % 9);
    const double v = S64[o] + add[t];
    S64[o] = v;
    if (S32) S32[o] = (float)v;
}

__global__ __launch_bounds__(256) void ext_rhs_kernel(int rows, const int* __restrict__ row_dof,
                                                      const int* __restrict__ row_ptr, const int* __restrict__ col_dof,
                                                      const double* __restrict__ kval, const double* __restrict__ fval,
                                                      const double* __restrict__ row_dt, const double* __restrict__ xn,
                                                      const double* __restrict__ xs, const double* __restrict__ vn,
                                                      int general, double div, double* __restrict__ b,
                                                      int div_by_dt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double dt = row_dt[r];
    double acc = 0.0;
    for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
        const int j = col_dof[p];
        const double w = general ? (xn[j] - xs[j]) - vn[j] * dt : -(vn[j] * dt);
        acc += kval[p] * w;
    }
    b[row_dof[r]] += (fval[r] + acc) / (div_by_dt ? dt : div);
}

void launch_ext_operator(int count, const int* blk, const double* add, double* S64, float* S32, hipStream_t s) {
    if (count <= 0) return;
    launch_k(ext_operator_kernel, dim3((unsigned)((9 * (int64_t)count + 255) / 256)), dim3(256), 0, s, count, blk, add,
             S64, S32);
}

void launch_ext_rhs(int rows, const int* row_dof, const int* row_ptr, const int* col_dof, const double* kval,
                    const double* fval, const double* row_dt, const double* xn, const double* xs, const double* vn,
                    int general, double div, double* b, hipStream_t s, int div_by_dt) {
    if (rows <= 0) return;
    launch_k(ext_rhs_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, rows, row_dof, row_ptr, col_dof, kval, fval,
             row_dt, xn, xs, vn, general, div, b, div_by_dt);
}

void launch_rhs(int num_nodes, const int* node_ptr, const double* stage, const double* fg, const double* fext,
                const uint8_t* bcmask, double* b, hipStream_t s, double div, const double* damp_d,
                const double* damp_v, double damp_alpha, const int* node_inst, const double* inst_div) {
    if (num_nodes == 0) return;
    launch_k(rhs_kernel, xcd_grid(3 * (int64_t)num_nodes, 256), dim3(256), 0, s, num_nodes, node_ptr, stage, fg,
             fext, bcmask, b, div, damp_d, damp_v, damp_alpha, node_inst, inst_div);
}

// ------------------------------------------------------------------------------------------
// 3x3-block SpMV core: a group of G lanes per block row; lane l takes blocks rowptr[I] + l, + G, ...
// (each block is 9 contiguous values, so the group reads one contiguous span of the row); the row sum
// is a butterfly over the group, so every lane of the group ends with the full sum.
// fetch(J, x0, x1, x2) supplies the input vector at block column J; epi(I, lane, y0, y1, y2) consumes
// row I on every lane of its group.
// ------------------------------------------------------------------------------------------
template <typename TM, typename TX, int G, typename Fetch, typename Epi>
__device__ __forceinline__ void group_spmv(const Operator& op, const TM* __restrict__ S, Fetch fetch, Epi epi) {
    const int lane = threadIdx.x & (G - 1);
    const int gpb = blockDim.x / G;                 // groups per block
    const int gib = threadIdx.x / G;                // group in block
    // XCD-aware rows: workgroups are dispatched round-robin over the 8 XCDs, so logical block
    // L = (b % 8) * (grid / 8) + b / 8 gives each XCD one contiguous slab of rows (and of the columns
    // its gathers touch), keeping those gathers in its own L2.
    const int nb = gridDim.x;
    const int L = (nb % kXcds == 0) ? (blockIdx.x % kXcds) * (nb / kXcds) + blockIdx.x / kXcds : blockIdx.x;
    const int per = (op.num_nodes + nb - 1) / nb;
    const int r0 = L * per;
    const int r1 = min(op.num_nodes, r0 + per);
    for (int I = r0 + gib; I < r1; I += gpb) {
        const int k0 = op.rowptr[I], k1 = op.rowptr[I + 1];
        TX y0 = 0, y1 = 0, y2 = 0;
        int jn = k0 + lane < k1 ? op.colidx[k0 + lane] : 0;
        for (int k = k0 + lane; k < k1; k += G) {
            const int J = jn;
            if (k + G < k1) jn = op.colidx[k + G];  // the next block's column goes out before this block's gathers
            TX p0, p1, p2;
            fetch(J, p0, p1, p2);
            const TM* B = S + 9 * (int64_t)k;
            y0 += (TX)B[0] * p0 + (TX)B[1] * p1 + (TX)B[2] * p2;
            y1 += (TX)B[3] * p0 + (TX)B[4] * p1 + (TX)B[5] * p2;
            y2 += (TX)B[6] * p0 + (TX)B[7] * p1 + (TX)B[8] * p2;
        }
        y0 = group_sum<G>(y0);  // the whole group is active here (I is uniform over the group)
        y1 = group_sum<G>(y1);
        y2 = group_sum<G>(y2);
        epi(I, lane, y0, y1, y2);
    }
}

// ------------------------------------------------------------------------------------------
// Conjugate gradient (Eigen 3.3 internal::conjugate_gradient recurrence, x0 = 0), two kernels per
// iteration on a fixed grid of kCgBlocks x kCgThreads (full occupancy). Each kernel publishes its
// block sums; the next kernel reduces them in block order in every block (deterministic, no atomics)
// and block 0 records the scalars that later kernels need.
// Partial slots: [0, B) |r|^2 or |b|^2, [B, 2B) r.z, [2B, 3B) p.q, [3B, 4B) FP64 residual.
// ------------------------------------------------------------------------------------------
// x = 0, r = (TV) b, dinv = 1/diag (1 where the diagonal is 0, DiagonalPreconditioner), z = dinv r;
// partial |r|^2 and r.z. Skipped when skip && skip->done (FP64 refinement already converged).
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_init_kernel(Operator op, const TV* __restrict__ S,
                                                             const double* __restrict__ b, TV* __restrict__ x,
                                                             TV* __restrict__ r, TV* __restrict__ z,
                                                             TV* __restrict__ dinv, double* __restrict__ partial,
                                                             const CgScalars* __restrict__ skip) {
    __shared__ double lds[64];
    if (skip && skip->done) return;
    const int n = 3 * op.num_nodes;
    double bb = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int I = i / 3, c = i - 3 * I;
        const int db = op.diagblk[I];  // -1: a node in no element (empty row)
        const TV s = db >= 0 ? S[9 * (int64_t)db + 4 * c] : TV(0);
        const TV d = (s != TV(0)) ? TV(1) / s : TV(1);
        const TV bi = (TV)b[i];
        x[i] = TV(0);
        r[i] = bi;
        dinv[i] = d;
        const TV zi = d * bi;
        z[i] = zi;
        bb += (double)bi * (double)bi;
        rz += (double)bi * (double)zi;
    }
    block_sum2(bb, rz, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bb;
        partial[kCgBlocks + blockIdx.x] = rz;
    }
}

// FP64 start from the guess in x (Eigen's solveWithGuess): r = b - S x, dinv as cg_init_kernel, z = dinv r;
// partial |b|^2, r.z and |r|^2 (slots 0, 1, 2 of the partials).
__global__ __launch_bounds__(kCgThreads) void cg_init_guess_kernel(Operator op, const double* __restrict__ S,
                                                                   const double* __restrict__ b,
                                                                   const double* __restrict__ x, double* __restrict__ r,
                                                                   double* __restrict__ z, double* __restrict__ dinv,
                                                                   double* __restrict__ partial) {
    __shared__ double lds[64];
    double bb = 0.0, rz = 0.0, rr = 0.0;
    group_spmv<double, double, kRowGroup>(
        op, S, [&](int J, double& p0, double& p1, double& p2) {
            p0 = x[3 * J];
            p1 = x[3 * J + 1];
            p2 = x[3 * J + 2];
        },
        [&](int I, int lane, double y0, double y1, double y2) {
            if (lane < 3) {
                const int i = 3 * I + lane;
                const int db = op.diagblk[I];
                const double sd = db >= 0 ? S[9 * (int64_t)db + 4 * lane] : 0.0;
                const double d = sd != 0.0 ? 1.0 / sd : 1.0;
                const double bi = b[i];
                const double ri = bi - (lane == 0 ? y0 : (lane == 1 ? y1 : y2));
                const double zi = d * ri;
                r[i] = ri;
                z[i] = zi;
                dinv[i] = d;
                bb += bi * bi;
                rz += ri * zi;
                rr += ri * ri;
            }
        });
    block_sum2(bb, rz, lds);  // ends with a barrier
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bb;
        partial[kCgBlocks + blockIdx.x] = rz;
    }
    __syncthreads();
    const double srr = block_sum(rr, lds);
    if (threadIdx.x == 0) partial[2 * kCgBlocks + blockIdx.x] = srr;
}

// Scalars of a solve started from a guess: threshold max(tol^2 |b|^2, realmin); Eigen's exit before the first
// iteration when |r|^2 is already below it (err = |r| / |b|), or when b = 0 (x = 0).
__global__ __launch_bounds__(kCgThreads) void cg_init2_guess_kernel(const double* __restrict__ partial, CgScalars* sc,
                                                                    double tol, double* __restrict__ x, int n) {
    __shared__ double lds[64];
    __shared__ int exit_zero;
    double bb, rz, rr;
    reduce_partials3(partial, partial + kCgBlocks, partial + 2 * kCgBlocks, bb, rz, rr, lds);
    if (threadIdx.x == 0) {
        sc->bb = bb;
        sc->rz[0] = rz;
        sc->rz[1] = 0.0;
        sc->threshold = fmax(tol * tol * bb, DBL_MIN);
        sc->rr = rr;
        sc->iters = 0;
        sc->err = 0.0;
        sc->converged = 0;
        sc->done = 0;
        exit_zero = 0;
        if (bb == 0.0) {
            sc->done = 1;
            sc->converged = 1;
            exit_zero = 1;
        } else if (rr < sc->threshold) {
            sc->done = 1;
            sc->converged = 1;
            sc->err = sqrt(rr / bb);
        }
    }
    __syncthreads();
    if (exit_zero)
        for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = 0.0;
}

// |b|^2 and r.z -> threshold max(tol^2 |b|^2, realmin) and Eigen's early exits. An inner solve of the FP64
// refinement (skip = the outer scalars) with floor_tol > 0 stops at the larger of that and floor_tol |b_outer|:
// its solution only has to bring the outer residual below the outer tolerance.
__global__ __launch_bounds__(kCgThreads) void cg_init2_kernel(const double* __restrict__ partial, CgScalars* sc,
                                                              const CgScalars* __restrict__ skip, double tol,
                                                              double floor_tol) {
    __shared__ double lds[64];
    if (skip && skip->done) return;
    double bb, rz;
    reduce_partials2(partial, partial + kCgBlocks, bb, rz, lds);
    if (threadIdx.x == 0) {
        double th = tol * tol * bb;
        if (skip && floor_tol > 0.0) th = fmax(th, floor_tol * floor_tol * skip->bb);
        sc->bb = bb;
        sc->rz[0] = rz;
        sc->rz[1] = 0.0;
        sc->threshold = fmax(th, DBL_MIN);
        sc->rr = bb;
        sc->iters = 0;
        sc->err = 0.0;
        sc->converged = 0;
        sc->done = 0;
        if (bb == 0.0) {
            sc->done = 1;
            sc->converged = 1;
        } else if (bb < sc->threshold) {
            sc->done = 1;
            sc->converged = 1;
            sc->err = 1.0;
        }
    }
}

// FP32 one-kernel iteration: x = 0, r = (float) b, the 3x3 block-diagonal preconditioner
// B = blockdiag(S)^-1 (inverted in FP64 from the node's diagonal block; a node whose block is missing or
// numerically singular falls back to the diagonal rule, 1 / S_ii or 1 where S_ii = 0), z = B r; partial
// |r|^2 and r.z as cg_init_kernel. One thread per node; skipped when skip && skip->done.
__global__ __launch_bounds__(kCgThreads) void cg_init_block_kernel(Operator op, const float* __restrict__ S,
                                                                   const double* __restrict__ b, float* __restrict__ x,
                                                                   float* __restrict__ r, float* __restrict__ z,
                                                                   float* __restrict__ binv,
                                                                   double* __restrict__ partial,
                                                                   const CgScalars* __restrict__ skip) {
    __shared__ double lds[64];
    if (skip && skip->done) return;
    double bb = 0.0, rz = 0.0;
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < op.num_nodes; I += gridDim.x * blockDim.x) {
        float Bf[9];
        block_jacobi_inverse(S, op.diagblk[I], Bf);  // diagblk -1: a node in no element (empty row)
        for (int k = 0; k < 9; ++k) binv[9 * (int64_t)I + k] = Bf[k];
        float rf[3];
        for (int c = 0; c < 3; ++c) {
            rf[c] = (float)b[3 * (int64_t)I + c];
            x[3 * (int64_t)I + c] = 0.f;
            r[3 * (int64_t)I + c] = rf[c];
        }
        for (int c = 0; c < 3; ++c) {
            const float zc = (Bf[3 * c] * rf[0] + Bf[3 * c + 1] * rf[1]) + Bf[3 * c + 2] * rf[2];
            z[3 * (int64_t)I + c] = zc;
            bb += (double)rf[c] * (double)rf[c];
            rz += (double)rf[c] * (double)zc;
        }
    }
    block_sum2(bb, rz, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bb;
        partial[kCgBlocks + blockIdx.x] = rz;
    }
}

// Iteration k, part 1. k > 0: |r|^2 of iteration k-1 below the threshold ends the solve with
// iters = k - 1 (Eigen breaks before incrementing), k == max_iter ends it unconverged; otherwise
// beta = rz_k / rz_{k-1}. Then p_k = z + beta p_{k-1}, formed on the fly for every column the row group
// needs and stored for its own row; q = S p_k; partial p_k . q.
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_apply_kernel(Operator op, const TV* __restrict__ S, int k,
                                                              int max_iter, const TV* __restrict__ z,
                                                              const TV* __restrict__ p_old, TV* __restrict__ p_new,
                                                              TV* __restrict__ q, double* __restrict__ partial,
                                                              CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    double beta = 0.0;
    if (k > 0) {
        double rr, rzk;
        reduce_partials2(partial, partial + kCgBlocks, rr, rzk, lds);
        const bool first = blockIdx.x == 0 && threadIdx.x == 0;
        if (rr < sc->threshold || k >= max_iter) {
            if (first) {
                const bool conv = rr < sc->threshold;
                sc->rr = rr;
                sc->err = sqrt(rr / sc->bb);
                sc->iters = conv ? k - 1 : k;
                sc->converged = conv ? 1 : 0;
                sc->done = 1;
            }
            return;
        }
        beta = rzk / sc->rz[(k - 1) & 1];
        if (first) {
            sc->rz[k & 1] = rzk;
            sc->rr = rr;
        }
    }
    const TV bt = (TV)beta;
    const bool k0 = (k == 0);
    double acc = 0.0;
    group_spmv<TV, TV, kRowGroup>(
        op, S,
        [&](int J, TV& a0, TV& a1, TV& a2) {
            if (k0) {
                a0 = z[3 * J];
                a1 = z[3 * J + 1];
                a2 = z[3 * J + 2];
            } else {
                a0 = z[3 * J] + bt * p_old[3 * J];
                a1 = z[3 * J + 1] + bt * p_old[3 * J + 1];
                a2 = z[3 * J + 2] + bt * p_old[3 * J + 2];
            }
        },
        [&](int I, int lane, TV y0, TV y1, TV y2) {
            if (lane < 3) {
                const int i = 3 * I + lane;
                const TV y = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
                const TV a = k0 ? z[i] : z[i] + bt * p_old[i];
                p_new[i] = a;
                q[i] = y;
                acc += (double)a * (double)y;
            }
        });
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[2 * kCgBlocks + blockIdx.x] = s;
}

// Iteration k, part 2: alpha = rz_k / (p . q); x += alpha p; r -= alpha q; z = dinv r; partials.
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_vector_kernel(int n, int k, TV* __restrict__ x, TV* __restrict__ r,
                                                               TV* __restrict__ z, const TV* __restrict__ p,
                                                               const TV* __restrict__ q,
                                                               const TV* __restrict__ dinv,
                                                               double* __restrict__ partial,
                                                               const CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    const double pq = reduce_partials(partial + 2 * kCgBlocks, lds);
    const TV a = (TV)(sc->rz[k & 1] / pq);
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        const TV ri = r[i] - a * q[i];
        r[i] = ri;
        const TV zi = dinv[i] * ri;
        z[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    block_sum2(rr, rz, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = rr;
        partial[kCgBlocks + blockIdx.x] = rz;
    }
}

// ------------------------------------------------------------------------------------------
// Fused CG iteration: one kernel per iteration of the same recurrence (Eigen's ConjugateGradient with
// the Jacobi preconditioner). Kernel k:
//   1. reduces the 7 block sums of kernel k - 1 over p.q and the products of r, z, q, w = D q
//      (D = dinv): alpha = (r.z) / (p.q) with r.z computed directly, as Eigen's absNew;
//      |r_k|^2 = r.r - 2 alpha r.q + alpha^2 q.q and r_k.z_k = r.z - 2 alpha q.z + alpha^2 q.w
//      (exact expansions of r_k = r - alpha q, z_k = z - alpha w);
//   2. applies the stopping test of iteration k - 1 (iters = k - 1 when converged, k at the cap), and
//      then only the x update of that iteration;
//   3. otherwise beta = (r_k.z_k) / (r.z); for its own rows r_k, z_k, x += alpha p, p_k = z_k + beta p;
//      p_k at every column the SpMV gathers is formed on the fly from (z, w, p); q_k = S p_k, w_k = D q_k;
//      block sums of p_k.q_k, r_k.r_k, r_k.q_k, q_k.q_k, r_k.z_k, q_k.z_k, q_k.w_k.
// r, z, w, p, q and the block sums are double-buffered: kernel k reads the buffers of parity (k + 1) & 1
// and writes those of parity k & 1 (a workgroup that finishes early must not overwrite sums that others
// have yet to read).
// The init kernel leaves r_0, z_0 in the odd buffers; kernel 0 copies them and forms p_0 = z_0.
// ------------------------------------------------------------------------------------------
static_assert(kCgThreads / 64 <= 16 && kFusedSums <= 64, "block_sums7 keeps 16 wave slots per sum");

// Block sums of the kFusedSums values: DPP wave totals into LDS, then lane j of wave 0 adds row j in wave
// order and writes it to out[j * kCgBlocks + blockIdx.x]. Every thread of the block calls it.
__device__ __forceinline__ void block_sums7(const double* v, double* lds, double* __restrict__ out) {
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int j = 0; j < kFusedSums; ++j) {
        const double t = wave_total(v[j]);
        if ((threadIdx.x & 63) == 0) lds[16 * j + w] = t;
    }
    __syncthreads();
    if (threadIdx.x < kFusedSums) {
        double s = 0.0;
        for (int q = 0; q < nw; ++q) s += lds[16 * threadIdx.x + q];
        out[threadIdx.x * kCgBlocks + blockIdx.x] = s;
    }
}

// The 7 sums of the nb partials of the last launch (nb = its grid size, at most kCgBlocks; rows of kCgBlocks):
// wave j reduces sum j in a fixed order; broadcast through LDS (lds[112 ..], clear of the rows block_sums7
// writes). A lane's kCgBlocks / 64 loads all go out before the first add (the partials were written on other
// XCDs, so each load misses this XCD's L2); the adds keep the order i = lane, lane + 64, ...
__device__ __forceinline__ void reduce7(const double* __restrict__ part, double* s, double* lds, int nb) {
    const int w = threadIdx.x >> 6;
    if (w < kFusedSums) {
        const int l = threadIdx.x & 63;
        const double* pw = part + w * kCgBlocks;
        double t[kCgBlocks / 64];
#pragma unroll
        for (int q = 0; q < kCgBlocks / 64; ++q) t[q] = pw[l + 64 * q < nb ? l + 64 * q : 0];
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < kCgBlocks / 64; ++q)
            if (l + 64 * q < nb) v += t[q];
        v = wave_total(v);
        if ((threadIdx.x & 63) == 0) lds[112 + w] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kFusedSums; ++j) s[j] = lds[112 + j];
}

template <typename TV>
__global__ __launch_bounds__(kCgThreads) __attribute__((amdgpu_waves_per_eu(sizeof(TV) == 4 ? 8 : 7))) void cg_fused_kernel(Operator op, const TV* __restrict__ S, int k,
                                                              int max_iter, TV* __restrict__ x, TV* r0, TV* r1,
                                                              TV* z0, TV* z1, TV* w0, TV* w1, TV* p0, TV* p1, TV* q0,
                                                              TV* q1, const TV* __restrict__ dinv,
                                                              double* __restrict__ part, CgScalars* __restrict__ sc) {
    __shared__ double lds[128];
    // Invariant: every workgroup of the launch that finds the solve converged (launch k, testing iteration
    // k - 1) must still apply that iteration's x += alpha p to its rows. Block 0 publishes completion as
    // done = 2 + k inside that same launch, and workgroups dispatched after its write would otherwise skip
    // the update; so only a flag from an earlier launch (or from another writer, done = 1) ends the kernel
    // here. The converging launch writes no block sums, so a late workgroup re-reads the same sums of
    // launch k - 1 and takes the same branch.
    const int done = sc->done;
    if (done != 0 && done != 2 + k) return;
    const bool odd = k & 1;
    const TV* ro = odd ? r0 : r1;
    const TV* zo = odd ? z0 : z1;
    const TV* wo = odd ? w0 : w1;
    const TV* po = odd ? p0 : p1;
    const TV* qo = odd ? q0 : q1;
    TV* rn = odd ? r1 : r0;
    TV* zn = odd ? z1 : z0;
    TV* wn = odd ? w1 : w0;
    TV* pn = odd ? p1 : p0;
    TV* qn = odd ? q1 : q0;
    // Rows of this workgroup (XCD-aware, as group_spmv) and the first row's own inputs: loaded before the
    // reduction of the block sums below, so their latency overlaps it.
    const int lane = threadIdx.x & (kRowGroup - 1);
    const int gpb = blockDim.x / kRowGroup, gib = threadIdx.x / kRowGroup;
    const int nb = gridDim.x;
    const int Lb = (nb % kXcds == 0) ? (blockIdx.x % kXcds) * (nb / kXcds) + blockIdx.x / kXcds : blockIdx.x;
    const int per = (op.num_nodes + nb - 1) / nb;
    const int row0 = Lb * per;
    const int row1 = min(op.num_nodes, row0 + per);
    const int cl = lane < 3 ? lane : 2;  // lanes 3.. of a group load lane 2's values (same line), unused
    int I = row0 + gib;
    int k0 = 0, k1 = 0;
    TV r_o = 0, q_o = 0, z_o = 0, w_o = 0, p_o = 0, x_o = 0, d_o = 0;
    // FP32 holds the row's own inputs across the SpMV (prefetched a row ahead); FP64 would need registers
    // beyond the 64 that keep two workgroups per CU, so it runs the plain row loop and reads them after.
    // FP32 (inner solves of the refinement) preconditions with the 3x3 block diagonal: dinv holds its
    // inverse, 9 per node, and lane c reads row c of its node's block (w = B q needs the row's q_0..2,
    // which every lane of the group has after the row sum). FP64 keeps Eigen's diagonal preconditioner.
    constexpr bool kPrefetch = sizeof(TV) == 4;
    TV d_o1 = 0, d_o2 = 0;
    int j_pf = 0;  // FP32: the column of the lane's first block of the row, fetched with the row's inputs
    auto fetch_row = [&](int row) {
        k0 = op.rowptr[row];
        k1 = op.rowptr[row + 1];
        if (k0 + lane < k1) j_pf = op.colidx[k0 + lane];
        const int i = 3 * row + cl;
        r_o = ro[i];
        q_o = qo[i];
        z_o = zo[i];
        w_o = wo[i];
        p_o = po[i];
        x_o = x[i];
        const TV* bq = dinv + 9 * (int64_t)row + 3 * cl;
        d_o = bq[0];
        d_o1 = bq[1];
        d_o2 = bq[2];
    };
    if (kPrefetch && I < row1) fetch_row(I);
    double alpha = 0.0, beta = 0.0;
    const double* part_old = part + (odd ? 0 : kFusedSums * kCgBlocks);
    double* part_new = part + (odd ? kFusedSums * kCgBlocks : 0);
    if (k > 0) {
        double sm[kFusedSums];
        reduce7(part_old, sm, lds, nb);  // the last launch had the same grid
        alpha = sm[4] / sm[0];
        const double rr = sm[1] - 2.0 * alpha * sm[2] + alpha * alpha * sm[3];
        const double rz = sm[4] - 2.0 * alpha * sm[5] + alpha * alpha * sm[6];
        const bool conv = rr < sc->threshold;
        if (conv || k >= max_iter) {
            const TV a = (TV)alpha;
            for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * op.num_nodes; i += gridDim.x * blockDim.x)
                x[i] += a * po[i];
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                sc->rr = rr;
                sc->err = sqrt(fmax(rr, 0.0) / sc->bb);
                sc->iters = conv ? k - 1 : k;
                sc->converged = conv ? 1 : 0;
                sc->done = 2 + k;
            }
            return;
        }
        beta = rz / sm[4];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->rz[k & 1] = rz;
            sc->rr = rr;
        }
    }
    const TV a = (TV)alpha, bt = (TV)beta;
    const bool k0it = k == 0;
    double acc[kFusedSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (!kPrefetch) {
        // FP64: the row loop of group_spmv, own-row inputs read after the SpMV
        group_spmv<TV, TV, kRowGroup>(
            op, S,
            [&](int J, TV& a0, TV& a1, TV& a2) {
                if (k0it) {
                    a0 = zo[3 * J];
                    a1 = zo[3 * J + 1];
                    a2 = zo[3 * J + 2];
                } else {
                    a0 = (zo[3 * J] - a * wo[3 * J]) + bt * po[3 * J];
                    a1 = (zo[3 * J + 1] - a * wo[3 * J + 1]) + bt * po[3 * J + 1];
                    a2 = (zo[3 * J + 2] - a * wo[3 * J + 2]) + bt * po[3 * J + 2];
                }
            },
            [&](int row, int ln, TV y0, TV y1, TV y2) {
                if (ln < 3) {
                    const int i = 3 * row + ln;
                    const TV y = ln == 0 ? y0 : (ln == 1 ? y1 : y2);
                    TV rk, zk, pk;
                    if (k0it) {
                        rk = ro[i];
                        zk = zo[i];
                        pk = zk;
                    } else {
                        rk = ro[i] - a * qo[i];
                        zk = zo[i] - a * wo[i];
                        pk = zk + bt * po[i];
                        x[i] += a * po[i];
                    }
                    const TV wk = dinv[i] * y;
                    rn[i] = rk;
                    zn[i] = zk;
                    pn[i] = pk;
                    qn[i] = y;
                    wn[i] = wk;
                    const double dr = (double)rk, dy = (double)y, dz = (double)zk;
                    acc[0] += (double)pk * dy;
                    acc[1] += dr * dr;
                    acc[2] += dr * dy;
                    acc[3] += dy * dy;
                    acc[4] += dr * dz;
                    acc[5] += dy * dz;
                    acc[6] += dy * (double)wk;
                }
            });
    } else {
        while (I < row1) {
            TV y0 = 0, y1 = 0, y2 = 0;
            int jn = j_pf;
            for (int kk = k0 + lane; kk < k1; kk += kRowGroup) {
                const int J = jn;
                if (kk + kRowGroup < k1) jn = op.colidx[kk + kRowGroup];  // the next block's column goes out early
                TV a0, a1, a2;
                if (k0it) {
                    a0 = zo[3 * J];
                    a1 = zo[3 * J + 1];
                    a2 = zo[3 * J + 2];
                } else {
                    a0 = (zo[3 * J] - a * wo[3 * J]) + bt * po[3 * J];
                    a1 = (zo[3 * J + 1] - a * wo[3 * J + 1]) + bt * po[3 * J + 1];
                    a2 = (zo[3 * J + 2] - a * wo[3 * J + 2]) + bt * po[3 * J + 2];
                }
                const TV* B = S + 9 * (int64_t)kk;
                y0 += B[0] * a0 + B[1] * a1 + B[2] * a2;
                y1 += B[3] * a0 + B[4] * a1 + B[5] * a2;
                y2 += B[6] * a0 + B[7] * a1 + B[8] * a2;
            }
            y0 = group_sum<kRowGroup>(y0);  // the whole group is active here (I is uniform over the group)
            y1 = group_sum<kRowGroup>(y1);
            y2 = group_sum<kRowGroup>(y2);
            const int row = I;
            const TV ro_i = r_o, qo_i = q_o, zo_i = z_o, wo_i = w_o, po_i = p_o, x_i = x_o;
            const TV wk = (d_o * y0 + d_o1 * y1) + d_o2 * y2;  // (B q)_c
            I += gpb;
            if (I < row1) fetch_row(I);  // the next row's inputs go out before this row's stores
            if (lane < 3) {
                const int i = 3 * row + lane;
                const TV y = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
                TV rk, zk, pk;
                if (k0it) {
                    rk = ro_i;
                    zk = zo_i;
                    pk = zk;
                } else {
                    rk = ro_i - a * qo_i;
                    zk = zo_i - a * wo_i;
                    pk = zk + bt * po_i;
                    x[i] = x_i + a * po_i;
                }
                rn[i] = rk;
                zn[i] = zk;
                pn[i] = pk;
                qn[i] = y;
                wn[i] = wk;
                const double dr = (double)rk, dy = (double)y, dz = (double)zk;
                acc[0] += (double)pk * dy;
                acc[1] += dr * dr;
                acc[2] += dr * dy;
                acc[3] += dy * dy;
                acc[4] += dr * dz;
                acc[5] += dy * dz;
                acc[6] += dy * (double)wk;
            }
        }
    }
    block_sums7(acc, lds, part_new);
}

// FP64 runs on half the grid: at the 40^3 block it takes 21.2 against 23.5 us per launch (fewer block sums to
// reduce at the start of every workgroup), while the FP32 iteration is as fast on either grid (14.2 us).
constexpr int kCgBlocksFused64 = kCgBlocks / 2;

void launch_cg_fused(const CgArgs& a, int k, int max_iter, hipStream_t s) {
    dim3 grid(a.fp32 ? kCgBlocks : kCgBlocksFused64), block(kCgThreads);
    if (a.fp32)
        launch_k((cg_fused_kernel<float>), grid, block, 0, s, a.op, (const float*)a.S, k, max_iter, (float*)a.x,
                 (float*)a.r, (float*)a.r1, (float*)a.z, (float*)a.z1, (float*)a.w[0], (float*)a.w[1],
                 (float*)a.p[0], (float*)a.p[1], (float*)a.q, (float*)a.q1, (const float*)a.binv, a.part7, a.sc);
    else
        launch_k((cg_fused_kernel<double>), grid, block, 0, s, a.op, (const double*)a.S, k, max_iter,
                 (double*)a.x, (double*)a.r, (double*)a.r1, (double*)a.z, (double*)a.z1, (double*)a.w[0],
                 (double*)a.w[1], (double*)a.p[0], (double*)a.p[1], (double*)a.q, (double*)a.q1,
                 (const double*)a.dinv, a.part7, a.sc);
}

void launch_cg_init(const CgArgs& a, const double* b, const CgScalars* skip, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    if (a.fp32)
        launch_k((cg_init_kernel<float>), grid, block, 0, s, a.op, (const float*)a.S, b, (float*)a.x,
                           (float*)a.r, (float*)a.z, (float*)a.dinv, a.partial, skip);
    else
        launch_k((cg_init_kernel<double>), grid, block, 0, s, a.op, (const double*)a.S, b, (double*)a.x,
                           (double*)a.r, (double*)a.z, (double*)a.dinv, a.partial, skip);
    launch_k(cg_init2_kernel, dim3(1), block, 0, s, a.partial, a.sc, skip, a.tol, a.floor_tol);
}

void launch_cg_init_guess(const CgArgs& a, const double* b, hipStream_t s) {
    launch_k(cg_init_guess_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a.op, (const double*)a.S, b,
             (const double*)a.x, (double*)a.r, (double*)a.z, (double*)a.dinv, a.partial);
    launch_k(cg_init2_guess_kernel, dim3(1), dim3(kCgThreads), 0, s, a.partial, a.sc, a.tol, (double*)a.x, a.n);
}

void launch_cg_init_block(const CgArgs& a, const double* b, const CgScalars* skip, hipStream_t s) {
    launch_k(cg_init_block_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a.op, (const float*)a.S, b, (float*)a.x,
             (float*)a.r, (float*)a.z, (float*)a.binv, a.partial, skip);
    launch_k(cg_init2_kernel, dim3(1), dim3(kCgThreads), 0, s, a.partial, a.sc, skip, a.tol, a.floor_tol);
}

void launch_cg_apply(const CgArgs& a, int k, int max_iter, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    void* pn = a.p[k & 1];
    const void* po = a.p[(k + 1) & 1];
    if (a.fp32)
        launch_k((cg_apply_kernel<float>), grid, block, 0, s, a.op, (const float*)a.S, k, max_iter,
                           (const float*)a.z, (const float*)po, (float*)pn, (float*)a.q, a.partial, a.sc);
    else
        launch_k((cg_apply_kernel<double>), grid, block, 0, s, a.op, (const double*)a.S, k, max_iter,
                           (const double*)a.z, (const double*)po, (double*)pn, (double*)a.q, a.partial, a.sc);
}

void launch_cg_init2(const double* partial, CgScalars* sc, const CgScalars* skip, double tol, hipStream_t s,
                     double floor_tol) {
    launch_k(cg_init2_kernel, dim3(1), dim3(kCgThreads), 0, s, partial, sc, skip, tol, floor_tol);
}

void launch_cg_vector_raw(bool fp32, int n, int k, void* x, void* r, void* z, const void* p, const void* q,
                          const void* dinv, double* partial, const CgScalars* sc, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    if (fp32)
        launch_k((cg_vector_kernel<float>), grid, block, 0, s, n, k, (float*)x, (float*)r, (float*)z,
                           (const float*)p, (const float*)q, (const float*)dinv, partial, sc);
    else
        launch_k((cg_vector_kernel<double>), grid, block, 0, s, n, k, (double*)x, (double*)r, (double*)z,
                           (const double*)p, (const double*)q, (const double*)dinv, partial, sc);
}

void launch_cg_vector(const CgArgs& a, int k, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    const void* p = a.p[k & 1];
    if (a.fp32)
        launch_k((cg_vector_kernel<float>), grid, block, 0, s, a.n, k, (float*)a.x, (float*)a.r,
                           (float*)a.z, (const float*)p, (const float*)a.q, (const float*)a.dinv, a.partial, a.sc);
    else
        launch_k((cg_vector_kernel<double>), grid, block, 0, s, a.n, k, (double*)a.x, (double*)a.r,
                           (double*)a.z, (const double*)p, (const double*)a.q, (const double*)a.dinv, a.partial,
                           a.sc);
}

// ------------------------------------------------------------------------------------------
// Newton-Raphson convergence test of OdeSolverEulerImplicit::solve (Math/OdeSolverEulerImplicit.cpp:101-112):
// ||dv - dv_prev||_inf. The grid kernel writes each workgroup's maximum to partial[blockIdx.x] and sets
// prev = dv for the next iteration (the loop leaves prev unused once the test passes); one workgroup then
// takes the maximum of the kCgBlocks partials into *out.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kCgThreads) void newton_variation_kernel(int64_t n, const double* __restrict__ dv,
                                                                      double* __restrict__ prev,
                                                                      double* __restrict__ partial) {
    __shared__ double lds[kCgThreads / 64];
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = dv[i];
        m = fmax(m, fabs(d - prev[i]));
        prev[i] = d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, lds[w]);
        partial[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void max_partials_kernel(const double* __restrict__ partial, int count,
                                                           double* __restrict__ out) {
    __shared__ double lds[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) m = fmax(m, partial[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

void launch_newton_variation(int64_t n, const double* dv, double* prev, double* partial, double* out,
                             hipStream_t s) {
    launch_k(newton_variation_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, dv, prev, partial);
    launch_k(max_partials_kernel, dim3(1), dim3(256), 0, s, (const double*)partial, (int)kCgBlocks, out);
}

// ------------------------------------------------------------------------------------------
// FP64 iterative refinement around the FP32 CG: r64 = b - S64 dv in FP64 (the first sweep takes
// r64 = b) with partial |r64|^2; the check kernel then applies the outer test
// |r64|^2 < max(tol^2 |b|^2, realmin) and marks the inner solve done when it passes. Once the outer
// test has passed, later sweeps' kernels return at once (sweeps can be queued ahead of the result).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kCgThreads) void residual64_kernel(Operator op, const double* __restrict__ S,
                                                                const double* __restrict__ b,
                                                                const double* __restrict__ dv, int first,
                                                                double* __restrict__ r64,
                                                                double* __restrict__ partial,
                                                                const CgScalars* __restrict__ outer) {
    __shared__ double lds[64];
    if (!first && outer->done) return;
    double acc = 0.0;
    if (first) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * op.num_nodes; i += gridDim.x * blockDim.x) {
            const double e = b[i];
            r64[i] = e;
            acc += e * e;
        }
    } else {
        group_spmv<double, double, kRowGroup>(
            op, S,
            [&](int J, double& a0, double& a1, double& a2) {
                a0 = dv[3 * J];
                a1 = dv[3 * J + 1];
                a2 = dv[3 * J + 2];
            },
            [&](int I, int lane, double y0, double y1, double y2) {
                if (lane < 3) {
                    const int i = 3 * I + lane;
                    const double e = b[i] - (lane == 0 ? y0 : (lane == 1 ? y1 : y2));
                    r64[i] = e;
                    acc += e * e;
                }
            });
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[3 * kCgBlocks + blockIdx.x] = s;
}

__global__ __launch_bounds__(kCgThreads) void refine_check_kernel(const double* __restrict__ partial, int first,
                                                                  double tol, CgScalars* outer, CgScalars* inner) {
    __shared__ double lds[64];
    if (!first && outer->done) return;
    const double rr = reduce_partials(partial + 3 * kCgBlocks, lds);
    if (threadIdx.x == 0) {
        if (first) {
            outer->bb = rr;
            outer->threshold = fmax(tol * tol * rr, DBL_MIN);
            outer->iters = 0;
            outer->done = 0;
            outer->converged = 0;
        }
        outer->rr = rr;
        outer->err = outer->bb > 0 ? sqrt(rr / outer->bb) : 0.0;
        if (outer->bb == 0.0 || rr < outer->threshold) {
            outer->done = 1;
            outer->converged = 1;
            inner->done = 1;
        }
    }
}

// dv += x (FP32 correction) unless the outer test has passed; records the inner solve's iteration
// count in *iters_out (left alone when skipped).
__global__ __launch_bounds__(kCgThreads) void accumulate_kernel(int n, const float* __restrict__ xf,
                                                                double* __restrict__ dv,
                                                                const CgScalars* __restrict__ outer,
                                                                const CgScalars* __restrict__ inner,
                                                                int* __restrict__ iters_out) {
    if (outer->done) return;
    if (iters_out && blockIdx.x == 0 && threadIdx.x == 0) *iters_out = inner->iters;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dv[i] += (double)xf[i];
}

// FP32 refinement on the one-kernel CG path, sweep s in one launch: dv_s = dv_{s-1} + x_{s-1} (s > 0; the
// row's own values written to dv_out, the gathers form it on the fly, so no launch adds the correction in
// place), the FP64 residual r64 = b - S64 dv_s, and the start of sweep s's FP32 inner solve as
// cg_init_block_kernel has it: x = 0, r = (float) r64, z = B r with the 3x3 block-diagonal inverse B (formed
// at s == 0 from S32 and kept for the step's later sweeps). Partial |r64|^2 (slot 3), |r|^2 (slot 0), r.z
// (slot 1). Skipped once the outer test has passed.
__global__ __launch_bounds__(kCgThreads) void refine_sweep_kernel(
    Operator op, const double* __restrict__ S64, const float* __restrict__ S32, const double* __restrict__ b,
    const double* __restrict__ dv_in, const float* __restrict__ x_in, double* __restrict__ dv_out, int first,
    float* __restrict__ x_out, float* __restrict__ r, float* __restrict__ z, float* __restrict__ binv,
    double* __restrict__ partial, const CgScalars* __restrict__ outer) {
    __shared__ double lds[64];
    if (!first && outer->done) return;
    double rr = 0.0, bb = 0.0, rz = 0.0;
    // lane c of node I's row starts DOF c of the inner solve from the node's residual e[0..2]
    auto start = [&](int I, const double* e, const float* B, int c) {
        const float r0 = (float)e[0], r1 = (float)e[1], r2 = (float)e[2];
        const float rc = c == 0 ? r0 : (c == 1 ? r1 : r2);
        const float zc = (B[3 * c] * r0 + B[3 * c + 1] * r1) + B[3 * c + 2] * r2;
        const int64_t i = 3 * (int64_t)I + c;
        x_out[i] = 0.f;
        r[i] = rc;
        z[i] = zc;
        bb += (double)rc * (double)rc;
        rz += (double)rc * (double)zc;
    };
    if (first) {
        for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < op.num_nodes; I += gridDim.x * blockDim.x) {
            float Bf[9];
            block_jacobi_inverse(S32, op.diagblk[I], Bf);  // diagblk -1: a node in no element (empty row)
            for (int k = 0; k < 9; ++k) binv[9 * (int64_t)I + k] = Bf[k];
            double e[3];
            for (int c = 0; c < 3; ++c) {
                e[c] = b[3 * (int64_t)I + c];
                dv_out[3 * (int64_t)I + c] = 0.0;
                rr += e[c] * e[c];
            }
            for (int c = 0; c < 3; ++c) start(I, e, Bf, c);
        }
    } else {
        group_spmv<double, double, kRowGroup>(
            op, S64,
            [&](int J, double& a0, double& a1, double& a2) {
                a0 = dv_in[3 * J] + (double)x_in[3 * J];
                a1 = dv_in[3 * J + 1] + (double)x_in[3 * J + 1];
                a2 = dv_in[3 * J + 2] + (double)x_in[3 * J + 2];
            },
            [&](int I, int lane, double y0, double y1, double y2) {
                if (lane < 3) {
                    const int64_t i0 = 3 * (int64_t)I;
                    const double e[3] = {b[i0] - y0, b[i0 + 1] - y1, b[i0 + 2] - y2};
                    dv_out[i0 + lane] = dv_in[i0 + lane] + (double)x_in[i0 + lane];
                    rr += e[lane] * e[lane];
                    start(I, e, binv + 9 * (int64_t)I, lane);
                }
            });
    }
    const double srr = block_sum(rr, lds);  // ends with a barrier: lds is free for the pair below
    block_sum2(bb, rz, lds);
    if (threadIdx.x == 0) {
        partial[3 * kCgBlocks + blockIdx.x] = srr;
        partial[blockIdx.x] = bb;
        partial[kCgBlocks + blockIdx.x] = rz;
    }
}

// The outer test of sweep s and, unless it passes, the scalars of sweep s's inner solve (refine_check_kernel and
// cg_init2_kernel in one launch). Records the iteration count of sweep s - 1's inner solve (s > 0) and clears the
// records at s == 0.
__global__ __launch_bounds__(kCgThreads) void refine_check_init_kernel(const double* __restrict__ partial, int sweep,
                                                                       double tol, double inner_tol,
                                                                       CgScalars* outer, CgScalars* inner,
                                                                       int* __restrict__ refine_iters, int nrec) {
    __shared__ double lds[64];
    if (sweep > 0 && outer->done) return;
    double rr, bb, rz;
    reduce_partials3(partial + 3 * kCgBlocks, partial, partial + kCgBlocks, rr, bb, rz, lds);
    if (threadIdx.x == 0) {
        if (sweep == 0) {
            for (int s = 0; s < nrec; ++s) refine_iters[s] = -1;
            outer->bb = rr;
            outer->threshold = fmax(tol * tol * rr, DBL_MIN);
            outer->iters = 0;
            outer->done = 0;
            outer->converged = 0;
        } else {
            refine_iters[sweep - 1] = inner->iters;
        }
        outer->rr = rr;
        outer->err = outer->bb > 0 ? sqrt(rr / outer->bb) : 0.0;
        if (outer->bb == 0.0 || rr < outer->threshold) {
            outer->done = 1;
            outer->converged = 1;
            inner->done = 1;
        } else {
            inner->bb = bb;
            inner->rz[0] = rz;
            inner->rz[1] = 0.0;
            // the inner solve only has to bring the outer residual below the outer tolerance (half of it)
            const double fl = kInnerFloor * tol;
            inner->threshold = fmax(fmax(inner_tol * inner_tol * bb, fl * fl * outer->bb), DBL_MIN);
            inner->rr = bb;
            inner->iters = 0;
            inner->err = 0.0;
            inner->converged = 0;
            inner->done = 0;
            if (bb == 0.0) {
                inner->done = 1;
                inner->converged = 1;
            } else if (bb < inner->threshold) {
                inner->done = 1;
                inner->converged = 1;
                inner->err = 1.0;
            }
        }
    }
}

void launch_refine_sweep(const Operator& op, const double* S64, const float* S32, const double* b,
                         const double* dv_in, const float* x_in, double* dv_out, int first, float* x_out, float* r,
                         float* z, float* binv, double* partial, const CgScalars* outer, hipStream_t s) {
    launch_k(refine_sweep_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, op, S64, S32, b, dv_in, x_in, dv_out,
             first, x_out, r, z, binv, partial, outer);
}

void launch_refine_check_init(const double* partial, int sweep, double tol, double inner_tol, CgScalars* outer,
                              CgScalars* inner, int* refine_iters, int nrec, hipStream_t s) {
    launch_k(refine_check_init_kernel, dim3(1), dim3(kCgThreads), 0, s, partial, sweep, tol, inner_tol, outer, inner,
             refine_iters, nrec);
}

void launch_residual64(const Operator& op, const double* S64, const double* b, const double* dv, int first,
                       double* r64, double* partial, CgScalars* outer, CgScalars* inner, double tol,
                       hipStream_t s) {
    launch_k(residual64_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, op, S64, b, dv, first, r64,
                       partial, (const CgScalars*)outer);
    launch_k(refine_check_kernel, dim3(1), dim3(kCgThreads), 0, s, partial, first, tol, outer, inner);
}

void launch_refine_accumulate(int n, const float* xf, double* dv, const CgScalars* outer, const CgScalars* inner,
                              int* iters_out, hipStream_t s) {
    launch_k(accumulate_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, xf, dv, outer, inner, iters_out);
}

// ------------------------------------------------------------------------------------------
// State update.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void update_kernel(int n, const double* __restrict__ dv, const double* xs, double* x,
                                                     double* v, double dt_all, int* invalid,
                                                     const int* __restrict__ node_inst,
                                                     const double* __restrict__ inst_dt,
                                                     const int* __restrict__ inst_active, int* inst_invalid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = node_inst ? node_inst[i / 3] : 0;
    if (inst_active && !inst_active[inst]) return;
    const double dt = node_inst ? inst_dt[inst] : dt_all;
    double vi = v[i] + dv[i];
    double xi = xs[i] + dt * vi;
    v[i] = vi;
    x[i] = xi;
    if (!isfinite(vi) || !isfinite(xi)) atomicOr(inst_invalid ? &inst_invalid[inst] : invalid, 1);
}

void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s, const int* node_inst, const double* inst_dt, const int* inst_active,
                   int* inst_invalid) {
    if (n == 0) return;
    launch_k(update_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, dv, xs, x, v, dt, invalid, node_inst, inst_dt,
             inst_active, inst_invalid);
}

void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s) {
    launch_update(n, dv, xs, x, v, dt, invalid, s, nullptr, nullptr);
}

// Math/OdeSolverEulerExplicit.cpp:37-38 (mode 1), Math/OdeSolverStatic.cpp:37-38 (mode 2); mode 0 is the
// implicit and modified explicit update (Math/OdeSolverEulerImplicit.cpp:98-99,
// Math/OdeSolverEulerExplicitModified.cpp:38-39).
__global__ __launch_bounds__(256) void update_scheme_kernel(int n, int mode, const double* __restrict__ dv,
                                                            const double* xs, double* x, double* v, double dt_all,
                                                            int* invalid, const int* __restrict__ node_inst,
                                                            const double* __restrict__ inst_dt,
                                                            const int* __restrict__ inst_active, int* inst_invalid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = node_inst ? node_inst[i / 3] : 0;
    if (inst_active && !inst_active[inst]) return;
    const double dt = node_inst ? inst_dt[inst] : dt_all;
    double vi, xi;
    if (mode == 1) {
        xi = xs[i] + dt * v[i];
        vi = v[i] + dv[i];
    } else if (mode == 2) {
        xi = xs[i] + dv[i];
        vi = 0.0;
    } else {
        vi = v[i] + dv[i];
        xi = xs[i] + dt * vi;
    }
    v[i] = vi;
    x[i] = xi;
    if (!isfinite(vi) || !isfinite(xi)) atomicOr(inst_invalid ? &inst_invalid[inst] : invalid, 1);
}

void launch_update_scheme(int n, int mode, const double* dv, const double* xs, double* x, double* v, double dt,
                          int* invalid, hipStream_t s, const int* node_inst, const double* inst_dt,
                          const int* inst_active, int* inst_invalid) {
    if (n == 0) return;
    launch_k(update_scheme_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, mode, dv, xs, x, v, dt, invalid,
             node_inst, inst_dt, inst_active, inst_invalid);
}

// Math/OdeSolverRungeKutta4.cpp:52-87: stage states y + k dt / 2 and y + k dt (dt per instance with node_inst).
__global__ __launch_bounds__(256) void rk4_stage_kernel(int n, const double* __restrict__ x,
                                                        const double* __restrict__ v, const double* __restrict__ kv,
                                                        const double* __restrict__ ka, double dt_all, double div,
                                                        double* __restrict__ xo, double* __restrict__ vo,
                                                        const int* __restrict__ node_inst,
                                                        const double* __restrict__ inst_dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = node_inst ? inst_dt[node_inst[i / 3]] : dt_all;
    xo[i] = x[i] + kv[i] * dt / div;
    vo[i] = v[i] + ka[i] * dt / div;
}

// Math/OdeSolverRungeKutta4.cpp:89-93 (k1v is the start velocity v)
__global__ __launch_bounds__(256) void rk4_final_kernel(int n, double* __restrict__ x, double* __restrict__ v,
                                                        const double* __restrict__ k2v, const double* __restrict__ k3v,
                                                        const double* __restrict__ k4v, const double* __restrict__ a1,
                                                        const double* __restrict__ a2, const double* __restrict__ a3,
                                                        const double* __restrict__ a4, double dt_all, int* invalid,
                                                        const int* __restrict__ node_inst,
                                                        const double* __restrict__ inst_dt,
                                                        const int* __restrict__ inst_active, int* inst_invalid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = node_inst ? node_inst[i / 3] : 0;
    if (inst_active && !inst_active[inst]) return;
    const double dt = node_inst ? inst_dt[inst] : dt_all;
    const double xi = x[i] + (v[i] + k4v[i] + 2.0 * (k2v[i] + k3v[i])) * dt / 6.0;
    const double vi = v[i] + (a1[i] + a4[i] + 2.0 * (a2[i] + a3[i])) * dt / 6.0;
    x[i] = xi;
    v[i] = vi;
    if (!isfinite(vi) || !isfinite(xi)) atomicOr(inst_invalid ? &inst_invalid[inst] : invalid, 1);
}

void launch_rk4_stage(int n, const double* x, const double* v, const double* kv, const double* ka, double dt,
                      double div, double* xo, double* vo, hipStream_t s, const int* node_inst, const double* inst_dt) {
    if (n == 0) return;
    launch_k(rk4_stage_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, x, v, kv, ka, dt, div, xo, vo, node_inst,
             inst_dt);
}

void launch_rk4_final(int n, double* x, double* v, const double* k2v, const double* k3v, const double* k4v,
                      const double* a1, const double* a2, const double* a3, const double* a4, double dt, int* invalid,
                      hipStream_t s, const int* node_inst, const double* inst_dt, const int* inst_active,
                      int* inst_invalid) {
    if (n == 0) return;
    launch_k(rk4_final_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, x, v, k2v, k3v, k4v, a1, a2, a3, a4, dt,
             invalid, node_inst, inst_dt, inst_active, inst_invalid);
}

// ------------------------------------------------------------------------------------------
// Element evaluation for tests: full R, K and f per element in FP64.
// ------------------------------------------------------------------------------------------
__global__ void eval_tet_kernel(int kind, int first, int count, const int4* conn, const double* x0all,
                                const double* xall, const double4* mat, double* Rout, double* Kout, double* fout) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int e = first + t;
    int4 id4 = conn[e];
    int ids[4] = {id4.x, id4.y, id4.z, id4.w};
    double x0[12], x[12];
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 3; ++c) {
            x0[3 * a + c] = x0all[3 * (size_t)ids[a] + c];
            x[3 * a + c] = xall[3 * (size_t)ids[a] + c];
        }
    double4 m = mat[e];
    double f[12];
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double g[4][3];
    double V = tet_rest_gradients(x0, g);
    CoroFactors<double> cf;
    if (kind == 1) {
        coro_factors<double>(x0, x, g, V, m.x, m.y, f, cf);
        for (int i = 0; i < 9; ++i) R[i] = cf.R[i];
    }
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double K[9];
            if (kind == 1) coro_block<double>(cf, a, b, K);
            else linear_block<double>(g, V, m.x, m.y, a, b, K);
            if (Kout)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Kout[(size_t)t * 144 + (3 * a + i) * 12 + 3 * b + j] = K[3 * i + j];
            if (kind != 1) {
                if (b == 0) f[3 * a] = f[3 * a + 1] = f[3 * a + 2] = 0.0;
                for (int i = 0; i < 3; ++i)
                    f[3 * a + i] -= K[3 * i] * (x[3 * b] - x0[3 * b]) + K[3 * i + 1] * (x[3 * b + 1] - x0[3 * b + 1]) +
                                    K[3 * i + 2] * (x[3 * b + 2] - x0[3 * b + 2]);
            }
        }
    if (Rout)
        for (int i = 0; i < 9; ++i) Rout[(size_t)t * 9 + i] = R[i];
    if (fout)
        for (int i = 0; i < 12; ++i) fout[(size_t)t * 12 + i] = f[i];
}

void launch_eval_elements(int kind, int first, int count, const int4* conn, const double* x0, const double* x,
                          const double4* mat, double* R, double* K, double* f, hipStream_t s) {
    if (count == 0) return;
    launch_k(eval_tet_kernel, dim3((count + 127) / 128), dim3(128), 0, s, kind, first, count, conn, x0, x,
                       mat, R, K, f);
}

}  // namespace femgpu
