// gfx950 kernels of the Fem3D implicit step: tetrahedron element kernels, deterministic gather
// assembly into a 3x3-block CSR operator, right-hand side, Jacobi-preconditioned CG and the
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

#include <cfloat>

namespace femgpu {

// ------------------------------------------------------------------------------------------
// Reductions (deterministic: fixed lane/wave order).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the block; result valid in thread 0. lds must hold blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    int nw = blockDim.x >> 6;
    double t = 0.0;
    if (threadIdx.x < 64) {
        t = (threadIdx.x < (unsigned)nw) ? lds[threadIdx.x] : 0.0;
        t = wave_sum(t);
    }
    __syncthreads();
    return t;
}

// Sum of kCgBlocks partials, broadcast to every thread of the block.
__device__ __forceinline__ double reduce_partials(const double* part, double* lds) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kCgBlocks; i += blockDim.x) v += part[i];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) lds[63] = v;
    __syncthreads();
    double r = lds[63];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------------------------
// Tetrahedron element kernels.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int pair_slot(int a, int b) { return a * (7 - a) / 2 + (b - a - 1); }

template <typename TK, bool COROT>
__global__ __launch_bounds__(256) void tet_element_kernel(TetElementArgs A) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.num_elements) return;
    int4 id4 = A.conn[e];
    int ids[4] = {id4.x, id4.y, id4.z, id4.w};
    double x0[12], xn[12], d1[12], d2[12];
    const double dt = A.dt;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            size_t i = 3 * (size_t)ids[a] + c;
            x0[3 * a + c] = A.x0[i];
            double xv = A.xn[i];
            double vv = A.vn[i];
            xn[3 * a + c] = xv;
            d1[3 * a + c] = (xv - A.xs[i]) - vv * dt;
            d2[3 * a + c] = vv - A.vs[i];
        }
    }
    double4 m = A.mat[e];
    const double lam = m.x, mu = m.y, rho = m.z;
    double f[12];
    double V;
    int flags = 0;
    CoroFactors<TK> cf;
    double g[4][3];
    V = tet_rest_gradients(x0, g);
    if (COROT) {
        flags = coro_factors<TK>(x0, xn, g, V, lam, mu, f, cf);
    } else {
        // f = -K0 (x - x0)  (Physics/Fem3DElementTetrahedron.cpp:77-90)
        for (int a = 0; a < 4; ++a) {
            double t[3] = {0.0, 0.0, 0.0};
            for (int b = 0; b < 4; ++b) {
                double Kb[9];
                linear_block<double>(g, V, lam, mu, a, b, Kb);
                double u[3] = {xn[3 * b] - x0[3 * b], xn[3 * b + 1] - x0[3 * b + 1], xn[3 * b + 2] - x0[3 * b + 2]};
                for (int i = 0; i < 3; ++i) t[i] += Kb[3 * i] * u[0] + Kb[3 * i + 1] * u[1] + Kb[3 * i + 2] * u[2];
            }
            f[3 * a] = -t[0];
            f[3 * a + 1] = -t[1];
            f[3 * a + 2] = -t[2];
        }
    }
    if (flags) {
        if (flags & ELEM_FLAG_DET_R) atomicAdd(&A.flag_counts[0], 1);
        if (flags & ELEM_FLAG_G_SINGULAR) atomicAdd(&A.flag_counts[1], 1);
        if (flags & ELEM_FLAG_POLAR_EIG) atomicAdd(&A.flag_counts[2], 1);
    }
    // Mass: M_e = R M0 R^T = rho V / 20 (2 on the diagonal blocks, 1 off) x I3.
    const double mc = rho * V / 20.0;
    const double idt = 1.0 / dt;
    TK* rec = reinterpret_cast<TK*>(A.rec) + (size_t)e * kTetRec;
    double* rr = A.rhs_rec + (size_t)e * 12;
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
        double ra[3] = {f[3 * a], f[3 * a + 1], f[3 * a + 2]};
        for (int b = 0; b < 4; ++b) {
            TK K[9];
            if (COROT) coro_block<TK>(cf, a, b, K);
            else linear_block<TK>(g, V, lam, mu, a, b, K);
            // rhs = f + K (xn - xs - vn dt) - M (vn - vs) / dt   (Math/OdeSolverEulerImplicit.cpp:157-158)
            const double mab = (a == b) ? 2.0 * mc : mc;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double kd = (double)K[3 * i] * d1[3 * b] + (double)K[3 * i + 1] * d1[3 * b + 1] +
                            (double)K[3 * i + 2] * d1[3 * b + 2];
                ra[i] += kd - (mab * d2[3 * b + i]) / dt;
            }
            if (!A.write_matrix) continue;
            // A_ab = M_ab * (1/dt) + K_ab * dt   (D = 0 for 3D elements, Physics/FemElement.cpp:131-145)
            const double mdt = mab * idt;
            if (a == b) {
                TK* o = rec + kTetPairSlots + 6 * a;
                o[0] = (TK)(mdt + (double)K[0] * dt);
                o[1] = (TK)((double)K[3] * dt);
                o[2] = (TK)(mdt + (double)K[4] * dt);
                o[3] = (TK)((double)K[6] * dt);
                o[4] = (TK)((double)K[7] * dt);
                o[5] = (TK)(mdt + (double)K[8] * dt);
            } else if (ids[a] > ids[b]) {
                int ps = (a < b) ? pair_slot(a, b) : pair_slot(b, a);
                TK* o = rec + 9 * ps;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        o[3 * i + j] = (TK)(((i == j) ? mdt : 0.0) + (double)K[3 * i + j] * dt);
            }
        }
        rr[3 * a] = ra[0];
        rr[3 * a + 1] = ra[1];
        rr[3 * a + 2] = ra[2];
    }
}

void launch_tet_elements(int kind, bool fp32, const TetElementArgs& a, hipStream_t s) {
    dim3 grid((a.num_elements + 255) / 256), block(256);
    if (a.num_elements == 0) return;
    if (kind == 1) {
        if (fp32) hipLaunchKernelGGL((tet_element_kernel<float, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((tet_element_kernel<double, true>), grid, block, 0, s, a);
    } else {
        if (fp32) hipLaunchKernelGGL((tet_element_kernel<float, false>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((tet_element_kernel<double, false>), grid, block, 0, s, a);
    }
}

// ------------------------------------------------------------------------------------------
// Deterministic gather assembly: one thread per lower block (I >= J) sums its contributions in a
// fixed (element) order, applies the boundary conditions (zero BC rows and columns, 1 on the BC
// diagonal, Math/OdeState.cpp:197-218) and writes the block and its transposed mirror, giving the
// lower-triangle self-adjoint view that Eigen's CG applies (Math/LinearSparseSolveAndInverse.h:122).
// ------------------------------------------------------------------------------------------
template <typename TK>
__global__ __launch_bounds__(256) void assemble_kernel(int num_lower, const int* __restrict__ lower_blk,
                                                       const int* __restrict__ mirror_blk,
                                                       const int* __restrict__ blk_row,
                                                       const int* __restrict__ blk_col,
                                                       const int* __restrict__ ctr_ptr,
                                                       const int* __restrict__ ctr_list, const TK* __restrict__ rec,
                                                       const uint8_t* __restrict__ bcmask, TK* __restrict__ S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_lower) return;
    int blk = lower_blk[t];
    int I = blk_row[blk];
    int J = blk_col[blk];
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0;
    int c0 = ctr_ptr[t], c1 = ctr_ptr[t + 1];
    if (I != J) {
        for (int c = c0; c < c1; ++c) {
            int code = ctr_list[c];
            const TK* r = rec + (size_t)(code >> 4) * kTetRec + 9 * (code & 15);
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] += (double)r[k];
        }
    } else {
        for (int c = c0; c < c1; ++c) {
            int code = ctr_list[c];
            const TK* r = rec + (size_t)(code >> 4) * kTetRec + kTetPairSlots + 6 * ((code & 15) - 6);
            acc[0] += (double)r[0];
            acc[3] += (double)r[1];
            acc[4] += (double)r[2];
            acc[6] += (double)r[3];
            acc[7] += (double)r[4];
            acc[8] += (double)r[5];
        }
        acc[1] = acc[3];
        acc[2] = acc[6];
        acc[5] = acc[7];
    }
    uint8_t br[3], bcol[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        br[k] = bcmask[3 * I + k];
        bcol[k] = bcmask[3 * J + k];
    }
    TK out[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double v = acc[3 * i + j];
            if (br[i] | bcol[j]) v = (I == J && i == j) ? 1.0 : 0.0;
            out[3 * i + j] = (TK)v;
        }
    TK* o = S + (size_t)blk * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = out[k];
    if (I != J) {
        TK* om = S + (size_t)mirror_blk[t] * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) om[3 * j + i] = out[3 * i + j];
    }
}

void launch_assemble(bool fp32, int num_lower, const int* lower_blk, const int* mirror_blk, const int* blk_row,
                     const int* blk_col, const int* ctr_ptr, const int* ctr_list, const void* rec,
                     const uint8_t* bcmask, void* S, hipStream_t s) {
    if (num_lower == 0) return;
    dim3 grid((num_lower + 255) / 256), block(256);
    if (fp32)
        hipLaunchKernelGGL((assemble_kernel<float>), grid, block, 0, s, num_lower, lower_blk, mirror_blk, blk_row,
                           blk_col, ctr_ptr, ctr_list, (const float*)rec, bcmask, (float*)S);
    else
        hipLaunchKernelGGL((assemble_kernel<double>), grid, block, 0, s, num_lower, lower_blk, mirror_blk, blk_row,
                           blk_col, ctr_ptr, ctr_list, (const double*)rec, bcmask, (double*)S);
}

// Right-hand side per node: element contributions in element order, gravity m_i g
// (Physics/FemRepresentation.cpp:641-662), external force; zero on BC DOFs (Math/OdeState.cpp:159-172).
__global__ __launch_bounds__(256) void rhs_kernel(int num_nodes, const int* __restrict__ node_ptr,
                                                  const int* __restrict__ node_list,
                                                  const double* __restrict__ rhs_rec, const double* __restrict__ fg,
                                                  const double* __restrict__ fext, const uint8_t* __restrict__ bcmask,
                                                  double* __restrict__ b) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= num_nodes) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = node_ptr[I]; c < node_ptr[I + 1]; ++c) {
        int code = node_list[c];
        const double* r = rhs_rec + (size_t)(code >> 2) * 12 + 3 * (code & 3);
        acc[0] += r[0];
        acc[1] += r[1];
        acc[2] += r[2];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        size_t i = 3 * (size_t)I + k;
        double v = acc[k] + fg[i];
        if (fext) v += fext[i];
        b[i] = bcmask[i] ? 0.0 : v;
    }
}

void launch_rhs(int num_nodes, const int* node_ptr, const int* node_list, const double* rhs_rec, const double* fg,
                const double* fext, const uint8_t* bcmask, double* b, hipStream_t s) {
    if (num_nodes == 0) return;
    hipLaunchKernelGGL(rhs_kernel, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, node_ptr, node_list,
                       rhs_rec, fg, fext, bcmask, b);
}

// ------------------------------------------------------------------------------------------
// Conjugate gradient (Eigen 3.3 internal::conjugate_gradient recurrence, x0 = 0).
// Kernels run on a fixed grid of kCgBlocks blocks; dot products are reduced per block into
// partials that the next kernel reduces in a fixed order (bit-reproducible runs).
// ------------------------------------------------------------------------------------------
template <typename TM, typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_init_kernel(int num_nodes, const int* __restrict__ diagblk,
                                                             const TM* __restrict__ S, const TV* __restrict__ b,
                                                             TV* __restrict__ x, TV* __restrict__ r,
                                                             TV* __restrict__ p, TV* __restrict__ dinv,
                                                             double* __restrict__ partial) {
    __shared__ double lds[64];
    int n = 3 * num_nodes;
    double bb = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int I = i / 3, c = i - 3 * I;
        TM s = S[(size_t)diagblk[I] * 9 + 4 * c];
        // DiagonalPreconditioner::factorize: 1/diag, 1 where the diagonal is zero
        TV d = (s != TM(0)) ? TV(1) / (TV)s : TV(1);
        TV bi = b[i];
        x[i] = TV(0);
        r[i] = bi;
        dinv[i] = d;
        TV pi = d * bi;
        p[i] = pi;
        bb += (double)bi * (double)bi;
        rz += (double)bi * (double)pi;
    }
    double s0 = block_sum(bb, lds);
    double s1 = block_sum(rz, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kCgBlocks + blockIdx.x] = s1;
    }
}

__global__ __launch_bounds__(kCgThreads) void cg_init2_kernel(const double* __restrict__ partial, CgScalars* sc,
                                                              double tol) {
    __shared__ double lds[64];
    double bb = reduce_partials(partial, lds);
    double rz = reduce_partials(partial + kCgBlocks, lds);
    if (threadIdx.x == 0) {
        sc->bb = bb;
        sc->rz[0] = rz;
        sc->rz[1] = 0.0;
        sc->threshold = fmax(tol * tol * bb, DBL_MIN);
        sc->iters = 0;
        sc->last_rr = bb;
        sc->done = 0;
        sc->converged = 0;
        sc->err = 0.0;
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

// q = S p (one group of G lanes per block row) and partial p . q.
template <typename TM, typename TV, int G>
__global__ __launch_bounds__(kCgThreads) void cg_spmv_kernel(int num_nodes, const int* __restrict__ rowptr,
                                                             const int* __restrict__ colidx,
                                                             const TM* __restrict__ S, const TV* __restrict__ p,
                                                             TV* __restrict__ q, double* __restrict__ partial,
                                                             const CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    const int lane = threadIdx.x & (G - 1);
    const int group = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int ngroups = gridDim.x * blockDim.x / G;
    double acc = 0.0;
    for (int I = group; I < num_nodes; I += ngroups) {
        TV y0 = 0, y1 = 0, y2 = 0;
        int k1 = rowptr[I + 1];
        for (int k = rowptr[I] + lane; k < k1; k += G) {
            int J = colidx[k];
            const TM* B = S + (size_t)k * 9;
            TV p0 = p[3 * J], p1 = p[3 * J + 1], p2 = p[3 * J + 2];
            y0 += (TV)B[0] * p0 + (TV)B[1] * p1 + (TV)B[2] * p2;
            y1 += (TV)B[3] * p0 + (TV)B[4] * p1 + (TV)B[5] * p2;
            y2 += (TV)B[6] * p0 + (TV)B[7] * p1 + (TV)B[8] * p2;
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            y0 += __shfl_xor(y0, o, G);
            y1 += __shfl_xor(y1, o, G);
            y2 += __shfl_xor(y2, o, G);
        }
        if (lane == 0) {
            q[3 * I] = y0;
            q[3 * I + 1] = y1;
            q[3 * I + 2] = y2;
            acc += (double)p[3 * I] * (double)y0 + (double)p[3 * I + 1] * (double)y1 + (double)p[3 * I + 2] * (double)y2;
        }
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[2 * kCgBlocks + blockIdx.x] = s;
}

// alpha = rz / pq; x += alpha p; r -= alpha q; partial r.r and r.z (z = dinv r).
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_update_kernel(int n, int k, TV* __restrict__ x, TV* __restrict__ r,
                                                               const TV* __restrict__ p, const TV* __restrict__ q,
                                                               const TV* __restrict__ dinv,
                                                               double* __restrict__ partial,
                                                               const CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    double pq = reduce_partials(partial + 2 * kCgBlocks, lds);
    double alpha = sc->rz[k & 1] / pq;
    TV a = (TV)alpha;
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        TV ri = r[i] - a * q[i];
        r[i] = ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)(dinv[i] * ri);
    }
    double s0 = block_sum(rr, lds);
    double s1 = block_sum(rz, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[kCgBlocks + blockIdx.x] = s1;
    }
}

// Convergence test on |r|^2, then beta = rz_new / rz; p = z + beta p.
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void cg_pupdate_kernel(int n, int k, int max_iter, const TV* __restrict__ r,
                                                                TV* __restrict__ p, const TV* __restrict__ dinv,
                                                                const double* __restrict__ partial,
                                                                CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    double rr = reduce_partials(partial, lds);
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr < sc->threshold) {
        if (first) {
            sc->last_rr = rr;
            sc->err = sqrt(rr / sc->bb);
            sc->iters = k;
            sc->converged = 1;
            sc->done = 1;
        }
        return;
    }
    double rzn = reduce_partials(partial + kCgBlocks, lds);
    double beta = rzn / sc->rz[k & 1];
    TV bt = (TV)beta;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = dinv[i] * r[i] + bt * p[i];
    if (first) {
        sc->rz[(k + 1) & 1] = rzn;
        sc->last_rr = rr;
        if (k + 1 >= max_iter) {
            sc->err = sqrt(rr / sc->bb);
            sc->iters = k + 1;
            sc->converged = 0;
            sc->done = 1;
        }
    }
}

template <typename TM, typename TV>
static void cg_iteration_t(const CgArgs& a, int k, int max_iter, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    hipLaunchKernelGGL((cg_spmv_kernel<TM, TV, 4>), grid, block, 0, s, a.num_nodes, a.rowptr, a.colidx,
                       (const TM*)a.S, (const TV*)a.p, (TV*)a.q, a.partial, a.sc);
    hipLaunchKernelGGL((cg_update_kernel<TV>), grid, block, 0, s, a.n, k, (TV*)a.x, (TV*)a.r, (const TV*)a.p,
                       (const TV*)a.q, (const TV*)a.dinv, a.partial, a.sc);
    hipLaunchKernelGGL((cg_pupdate_kernel<TV>), grid, block, 0, s, a.n, k, max_iter, (const TV*)a.r, (TV*)a.p,
                       (const TV*)a.dinv, a.partial, a.sc);
}

void launch_cg_iteration(const CgArgs& a, int k, int max_iter, hipStream_t s) {
    if (a.fp32) cg_iteration_t<float, float>(a, k, max_iter, s);
    else cg_iteration_t<double, double>(a, k, max_iter, s);
}

void launch_cg_init(const CgArgs& a, hipStream_t s) {
    dim3 grid(kCgBlocks), block(kCgThreads);
    if (a.fp32)
        hipLaunchKernelGGL((cg_init_kernel<float, float>), grid, block, 0, s, a.num_nodes, a.diagblk,
                           (const float*)a.S, (const float*)a.b, (float*)a.x, (float*)a.r, (float*)a.p,
                           (float*)a.dinv, a.partial);
    else
        hipLaunchKernelGGL((cg_init_kernel<double, double>), grid, block, 0, s, a.num_nodes, a.diagblk,
                           (const double*)a.S, (const double*)a.b, (double*)a.x, (double*)a.r, (double*)a.p,
                           (double*)a.dinv, a.partial);
    hipLaunchKernelGGL(cg_init2_kernel, dim3(1), block, 0, s, a.partial, a.sc, a.tol);
}

// ------------------------------------------------------------------------------------------
// FP64 iterative refinement around the FP32 CG.
// ------------------------------------------------------------------------------------------
// r64 = b - S dv with S in FP32 promoted to FP64; partial |r64|^2 into partial[3 * kCgBlocks + blk].
template <int G>
__global__ __launch_bounds__(kCgThreads) void residual64_kernel(int num_nodes, const int* __restrict__ rowptr,
                                                                const int* __restrict__ colidx,
                                                                const float* __restrict__ S,
                                                                const double* __restrict__ b,
                                                                const double* __restrict__ dv,
                                                                double* __restrict__ r64, double* __restrict__ partial) {
    __shared__ double lds[64];
    const int lane = threadIdx.x & (G - 1);
    const int group = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int ngroups = gridDim.x * blockDim.x / G;
    double acc = 0.0;
    for (int I = group; I < num_nodes; I += ngroups) {
        double y0 = 0, y1 = 0, y2 = 0;
        int k1 = rowptr[I + 1];
        for (int k = rowptr[I] + lane; k < k1; k += G) {
            int J = colidx[k];
            const float* B = S + (size_t)k * 9;
            double p0 = dv[3 * J], p1 = dv[3 * J + 1], p2 = dv[3 * J + 2];
            y0 += (double)B[0] * p0 + (double)B[1] * p1 + (double)B[2] * p2;
            y1 += (double)B[3] * p0 + (double)B[4] * p1 + (double)B[5] * p2;
            y2 += (double)B[6] * p0 + (double)B[7] * p1 + (double)B[8] * p2;
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            y0 += __shfl_xor(y0, o, G);
            y1 += __shfl_xor(y1, o, G);
            y2 += __shfl_xor(y2, o, G);
        }
        if (lane == 0) {
            double r0 = b[3 * I] - y0, r1 = b[3 * I + 1] - y1, r2 = b[3 * I + 2] - y2;
            r64[3 * I] = r0;
            r64[3 * I + 1] = r1;
            r64[3 * I + 2] = r2;
            acc += r0 * r0 + r1 * r1 + r2 * r2;
        }
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[3 * kCgBlocks + blockIdx.x] = s;
}

// |b|^2 partials for the first sweep.
__global__ __launch_bounds__(kCgThreads) void norm2_kernel(int n, const double* __restrict__ v,
                                                           double* __restrict__ partial) {
    __shared__ double lds[64];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += v[i] * v[i];
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[3 * kCgBlocks + blockIdx.x] = s;
}

// Outer test and the inner right-hand side. outer->bb holds |b|^2 after the first sweep; outer->done
// is set once |b - S dv|^2 < max(tol^2 |b|^2, realmin). The inner CG's scalars are marked done as well
// so that its kernels skip.
__global__ __launch_bounds__(kCgThreads) void refine_check_kernel(const double* __restrict__ partial, int first,
                                                                  double tol, CgScalars* outer, CgScalars* inner) {
    __shared__ double lds[64];
    double rr = reduce_partials(partial + 3 * kCgBlocks, lds);
    if (threadIdx.x == 0) {
        if (first) {
            outer->bb = rr;
            outer->threshold = fmax(tol * tol * rr, DBL_MIN);
            outer->iters = 0;
            outer->done = 0;
            outer->converged = 0;
        }
        outer->last_rr = rr;
        outer->err = outer->bb > 0 ? sqrt(rr / outer->bb) : 0.0;
        if (outer->bb == 0.0 || rr < outer->threshold) {
            outer->done = 1;
            outer->converged = 1;
            inner->done = 1;
        } else {
            inner->done = 0;
        }
    }
}

__global__ __launch_bounds__(kCgThreads) void to_float_kernel(int n, const double* __restrict__ src,
                                                              float* __restrict__ dst,
                                                              const CgScalars* __restrict__ outer) {
    if (outer->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = (float)src[i];
}

__global__ __launch_bounds__(kCgThreads) void accumulate_kernel(int n, const float* __restrict__ xf,
                                                                double* __restrict__ dv,
                                                                const CgScalars* __restrict__ outer) {
    if (outer->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dv[i] += (double)xf[i];
}

void launch_residual64(const CgArgs& a, const double* b, const double* dv, double* r64, hipStream_t s) {
    if (dv == nullptr) {
        hipLaunchKernelGGL(norm2_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a.n, b, a.partial);
        return;
    }
    hipLaunchKernelGGL((residual64_kernel<4>), dim3(kCgBlocks), dim3(kCgThreads), 0, s, a.num_nodes, a.rowptr,
                       a.colidx, (const float*)a.S, b, dv, r64, a.partial);
}

void launch_refine_prepare(const CgArgs& a, const double* r64, int first, double tol, CgScalars* outer,
                           hipStream_t s) {
    hipLaunchKernelGGL(refine_check_kernel, dim3(1), dim3(kCgThreads), 0, s, a.partial, first, tol, outer, a.sc);
    hipLaunchKernelGGL(to_float_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, a.n, r64, (float*)a.b, outer);
}

void launch_refine_accumulate(int n, const float* xf, double* dv, const CgScalars* outer, hipStream_t s) {
    hipLaunchKernelGGL(accumulate_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, xf, dv, outer);
}

// ------------------------------------------------------------------------------------------
// State update.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void update_kernel(int n, const double* __restrict__ dv, const double* xs, double* x,
                                                     double* v, double dt, int* invalid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double vi = v[i] + dv[i];
    double xi = xs[i] + dt * vi;
    v[i] = vi;
    x[i] = xi;
    if (!isfinite(vi) || !isfinite(xi)) atomicOr(invalid, 1);
}

void launch_update(int n, const double* dv, const double* xs, double* x, double* v, double dt, int* invalid,
                   hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(update_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, dv, xs, x, v, dt, invalid);
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
    hipLaunchKernelGGL(eval_tet_kernel, dim3((count + 127) / 128), dim3(128), 0, s, kind, first, count, conn, x0, x,
                       mat, R, K, f);
}

}  // namespace femgpu
