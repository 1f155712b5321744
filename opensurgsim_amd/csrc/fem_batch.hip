// Batched Eigen-CG over independent instances (disconnected meshes concatenated into one block-diagonal
// operator). The recurrence, stopping rule and iteration count of each instance are those of the
// single-instance solver (Math/LinearSparseSolveAndInverse.cpp:99-129 over Eigen's ConjugateGradient);
// only the launch layout differs: one workgroup per row chunk, chunks never straddle instances, each
// chunk publishes partial sums for its instance, and every consumer reduces the partials of its own
// instance in chunk order (deterministic, no atomics). Converged instances skip their chunks.
#include "fem_batch.hpp"
#include "fem_device.hpp"
#include "fem_reduce.hpp"

#include <cfloat>

namespace femgpu {

namespace {

// Sums of two partial arrays over chunks [c0, c1), by wave 0 in a fixed order, broadcast through LDS.
__device__ __forceinline__ void range_sum2(const double* pa, const double* pb, int c0, int c1, double& a,
                                           double& b, double* lds) {
    if (threadIdx.x < 64) {
        double va = 0.0, vb = 0.0;
        for (int i = c0 + (int)threadIdx.x; i < c1; i += 64) {
            va += pa[i];
            if (pb) vb += pb[i];
        }
        va = wave_total(va);
        vb = wave_total(vb);
        if (threadIdx.x == 0) {
            lds[62] = va;
            lds[63] = vb;
        }
    }
    __syncthreads();
    a = lds[62];
    b = lds[63];
    __syncthreads();
}

// Block rows [r0, r1) of y = S x with G lanes per row (butterfly sum within the group).
template <typename TM, typename TX, int G, typename Fetch, typename Epi>
__device__ __forceinline__ void chunk_spmv(const Operator& op, const TM* __restrict__ S, int r0, int r1, Fetch fetch,
                                           Epi epi) {
    const int lane = threadIdx.x & (G - 1);
    const int gpb = blockDim.x / G;
    const int gib = threadIdx.x / G;
    for (int I0 = r0; I0 < r1; I0 += gpb) {
        const int I = I0 + gib;
        const bool live = I < r1;  // every lane of a group takes the same branch
        TX y0 = 0, y1 = 0, y2 = 0;
        if (live) {
            const int k0 = op.rowptr[I], k1 = op.rowptr[I + 1];
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
        }
        y0 = group_sum<G>(y0);
        y1 = group_sum<G>(y1);
        y2 = group_sum<G>(y2);
        if (live) epi(I, lane, y0, y1, y2);
    }
}

template <typename TV>
__global__ __launch_bounds__(kBatchThreads) void bcg_init_kernel(BatchTables t, Operator op, const TV* __restrict__ S,
                                                                 const double* __restrict__ b, TV* __restrict__ x,
                                                                 TV* __restrict__ r, TV* __restrict__ z,
                                                                 TV* __restrict__ dinv, double* __restrict__ partial,
                                                                 const CgScalars* __restrict__ skip) {
    __shared__ double lds[64];
    const int c = blockIdx.x, inst = t.chunk_inst[c];
    const bool left = (skip && skip[inst].done) || (t.inst_skip && t.inst_skip[inst]);
    double bb = 0.0, rz = 0.0;
    if constexpr (sizeof(TV) == 4) {
        // FP32 inner solves of the refinement: 3x3 block-diagonal preconditioner, dinv = 9 floats per node
        if (!left) {
            for (int I = t.chunk_row0[c] + (int)threadIdx.x; I < t.chunk_row1[c]; I += blockDim.x) {
                float Bf[9];
                block_jacobi_inverse(S, op.diagblk[I], Bf);
                for (int k = 0; k < 9; ++k) dinv[9 * (int64_t)I + k] = Bf[k];
                float rf[3];
                for (int cc = 0; cc < 3; ++cc) {
                    rf[cc] = (float)b[3 * (int64_t)I + cc];
                    x[3 * (int64_t)I + cc] = 0.f;
                    r[3 * (int64_t)I + cc] = rf[cc];
                }
                for (int cc = 0; cc < 3; ++cc) {
                    const float zc = (Bf[3 * cc] * rf[0] + Bf[3 * cc + 1] * rf[1]) + Bf[3 * cc + 2] * rf[2];
                    z[3 * (int64_t)I + cc] = zc;
                    bb += (double)rf[cc] * (double)rf[cc];
                    rz += (double)rf[cc] * (double)zc;
                }
            }
        }
    } else if (!left) {
        const int d0 = 3 * t.chunk_row0[c], d1 = 3 * t.chunk_row1[c];
        for (int i = d0 + (int)threadIdx.x; i < d1; i += blockDim.x) {
            const int I = i / 3, cc = i - 3 * I;
            const int db = op.diagblk[I];  // -1: a node in no element (empty row)
            const TV s = db >= 0 ? S[9 * (int64_t)db + 4 * cc] : TV(0);
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
    }
    block_sum2(bb, rz, lds);
    if (threadIdx.x == 0) {
        partial[c] = bb;
        partial[t.num_chunks + c] = rz;
    }
}

// One workgroup per instance.
__global__ __launch_bounds__(kBatchThreads) void bcg_init2_kernel(BatchTables t, const double* __restrict__ partial,
                                                                  CgScalars* sc, const CgScalars* __restrict__ skip,
                                                                  double tol, int* kA, int* kV) {
    __shared__ double lds[64];
    const int inst = blockIdx.x;
    if (threadIdx.x == 0) {
        kA[inst] = 0;
        kV[inst] = 0;
    }
    if (skip && skip[inst].done) {  // no inner solve this sweep: no iterations of it either
        if (threadIdx.x == 0) {
            CgScalars s = {};
            s.done = 1;
            s.converged = 1;
            sc[inst] = s;
        }
        return;
    }
    double bb, rz;
    range_sum2(partial, partial + t.num_chunks, t.inst_chunk[inst], t.inst_chunk[inst + 1], bb, rz, lds);
    if (threadIdx.x == 0) {
        if (!t.inst_active[inst] || (t.inst_skip && t.inst_skip[inst])) {
            // deactivated instance: no solve (dv is zeroed by the init kernel); an instance this pass leaves
            // alone: no solve, its x untouched
            CgScalars s = {};
            s.done = 1;
            s.converged = 1;
            sc[inst] = s;
            return;
        }
        CgScalars s;
        s.bb = bb;
        s.rz[0] = rz;
        s.rz[1] = 0.0;
        s.threshold = fmax(tol * tol * bb, DBL_MIN);
        s.rr = bb;
        s.iters = 0;
        s.err = 0.0;
        s.converged = 0;
        s.done = 0;
        s.pad = 0;
        if (bb == 0.0) {
            s.done = 1;
            s.converged = 1;
        } else if (bb < s.threshold) {
            s.done = 1;
            s.converged = 1;
            s.err = 1.0;
        }
        sc[inst] = s;
    }
}

// Iteration k of every live instance, part 1 (k read from kA, recorded in kV): convergence test of
// k - 1, beta, p_k = z + beta p_{k-1} on the fly, q = S p_k, partial p_k . q.
template <typename TV, int G>
__global__ __launch_bounds__(kBatchThreads) void bcg_apply_kernel(BatchTables t, Operator op, const TV* __restrict__ S,
                                                                  const TV* __restrict__ z, TV* p0, TV* p1,
                                                                  TV* __restrict__ q, double* __restrict__ partial,
                                                                  CgScalars* sc, const int* __restrict__ kA,
                                                                  int* __restrict__ kV) {
    __shared__ double lds[64];
    const int c = t.live ? t.live[blockIdx.x] : blockIdx.x, inst = t.chunk_inst[c];
    if (sc[inst].done) return;
    const int k = kA[inst];
    const int K = t.num_chunks;
    const bool first = c == t.inst_chunk[inst] && threadIdx.x == 0;
    double beta = 0.0;
    if (k > 0) {
        double rr, rzk;
        range_sum2(partial, partial + K, t.inst_chunk[inst], t.inst_chunk[inst + 1], rr, rzk, lds);
        const int max_iter = t.inst_maxit[inst];
        if (rr < sc[inst].threshold || k >= max_iter) {
            if (first) {
                const bool conv = rr < sc[inst].threshold;
                sc[inst].rr = rr;
                sc[inst].err = sqrt(rr / sc[inst].bb);
                sc[inst].iters = conv ? k - 1 : k;
                sc[inst].converged = conv ? 1 : 0;
                sc[inst].done = 1;
            }
            return;
        }
        beta = rzk / sc[inst].rz[(k - 1) & 1];
        if (first) {
            sc[inst].rz[k & 1] = rzk;
            sc[inst].rr = rr;
        }
    }
    if (first) kV[inst] = k;
    TV* pn = (k & 1) ? p1 : p0;
    const TV* po = (k & 1) ? p0 : p1;
    const TV bt = (TV)beta;
    const bool k0 = k == 0;
    double acc = 0.0;
    chunk_spmv<TV, TV, G>(
        op, S, t.chunk_row0[c], t.chunk_row1[c],
        [&](int J, TV& a0, TV& a1, TV& a2) {
            if (k0) {
                a0 = z[3 * J];
                a1 = z[3 * J + 1];
                a2 = z[3 * J + 2];
            } else {
                a0 = z[3 * J] + bt * po[3 * J];
                a1 = z[3 * J + 1] + bt * po[3 * J + 1];
                a2 = z[3 * J + 2] + bt * po[3 * J + 2];
            }
        },
        [&](int I, int lane, TV y0, TV y1, TV y2) {
            if (lane < 3) {
                const int i = 3 * I + lane;
                const TV y = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
                const TV a = k0 ? z[i] : z[i] + bt * po[i];
                pn[i] = a;
                q[i] = y;
                acc += (double)a * (double)y;
            }
        });
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[2 * K + c] = s;
}

// Part 2 (k read from kV, k + 1 written to kA): alpha = rz_k / (p.q), x += alpha p, r -= alpha q,
// z = dinv r, partial |r|^2 and r.z.
template <typename TV>
__global__ __launch_bounds__(kBatchThreads) void bcg_vector_kernel(BatchTables t, TV* __restrict__ x,
                                                                   TV* __restrict__ r, TV* __restrict__ z,
                                                                   const TV* p0, const TV* p1,
                                                                   const TV* __restrict__ q,
                                                                   const TV* __restrict__ dinv,
                                                                   double* __restrict__ partial,
                                                                   const CgScalars* sc, const int* __restrict__ kV,
                                                                   int* __restrict__ kA) {
    __shared__ double lds[64];
    const int c = t.live ? t.live[blockIdx.x] : blockIdx.x, inst = t.chunk_inst[c];
    if (sc[inst].done) return;
    const int k = kV[inst];
    const int K = t.num_chunks;
    double pq, unused;
    range_sum2(partial + 2 * K, nullptr, t.inst_chunk[inst], t.inst_chunk[inst + 1], pq, unused, lds);
    const TV a = (TV)(sc[inst].rz[k & 1] / pq);
    const TV* p = (k & 1) ? p1 : p0;
    double rr = 0.0, rz = 0.0;
    if constexpr (sizeof(TV) == 4) {
        // block-diagonal preconditioner: a thread takes a whole node (z needs its three r components)
        for (int I = t.chunk_row0[c] + (int)threadIdx.x; I < t.chunk_row1[c]; I += blockDim.x) {
            TV rI[3];
            for (int cc = 0; cc < 3; ++cc) {
                const int64_t i = 3 * (int64_t)I + cc;
                x[i] += a * p[i];
                rI[cc] = r[i] - a * q[i];
                r[i] = rI[cc];
            }
            const TV* B = dinv + 9 * (int64_t)I;
            for (int cc = 0; cc < 3; ++cc) {
                const TV zc = (B[3 * cc] * rI[0] + B[3 * cc + 1] * rI[1]) + B[3 * cc + 2] * rI[2];
                z[3 * (int64_t)I + cc] = zc;
                rr += (double)rI[cc] * (double)rI[cc];
                rz += (double)rI[cc] * (double)zc;
            }
        }
    } else {
        const int d0 = 3 * t.chunk_row0[c], d1 = 3 * t.chunk_row1[c];
        for (int i = d0 + (int)threadIdx.x; i < d1; i += blockDim.x) {
            x[i] += a * p[i];
            const TV ri = r[i] - a * q[i];
            r[i] = ri;
            const TV zi = dinv[i] * ri;
            z[i] = zi;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)zi;
        }
    }
    block_sum2(rr, rz, lds);
    if (threadIdx.x == 0) {
        partial[c] = rr;
        partial[K + c] = rz;
        if (c == t.inst_chunk[inst]) kA[inst] = k + 1;
    }
}

template <int G>
__global__ __launch_bounds__(kBatchThreads) void bres64_kernel(BatchTables t, Operator op, const double* __restrict__ S,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ dv, int first,
                                                               double* __restrict__ r64, const CgScalars* outer,
                                                               double* __restrict__ partial) {
    __shared__ double lds[64];
    const int c = blockIdx.x, inst = t.chunk_inst[c];
    double acc = 0.0;
    if (first) {
        const int d0 = 3 * t.chunk_row0[c], d1 = 3 * t.chunk_row1[c];
        for (int i = d0 + (int)threadIdx.x; i < d1; i += blockDim.x) {
            const double e = b[i];
            r64[i] = e;
            acc += e * e;
        }
    } else if (!outer[inst].done) {
        chunk_spmv<double, double, G>(
            op, S, t.chunk_row0[c], t.chunk_row1[c],
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
    if (threadIdx.x == 0) partial[3 * t.num_chunks + c] = s;
}

// One workgroup per instance: the outer (FP64 residual) test of the refinement.
__global__ __launch_bounds__(kBatchThreads) void brefine_check_kernel(BatchTables t, const double* __restrict__ partial,
                                                                      int first, double tol, CgScalars* outer,
                                                                      CgScalars* inner) {
    __shared__ double lds[64];
    const int inst = blockIdx.x;
    if (!first && outer[inst].done) return;
    double rr, unused;
    range_sum2(partial + 3 * t.num_chunks, nullptr, t.inst_chunk[inst], t.inst_chunk[inst + 1], rr, unused, lds);
    if (threadIdx.x == 0) {
        CgScalars& o = outer[inst];
        if (first) {
            o.bb = rr;
            o.threshold = fmax(tol * tol * rr, DBL_MIN);
            o.iters = 0;
            o.done = 0;
            o.converged = 0;
        }
        o.rr = rr;
        o.err = o.bb > 0 ? sqrt(rr / o.bb) : 0.0;
        if (!t.inst_active[inst] || (t.inst_skip && t.inst_skip[inst]) || o.bb == 0.0 || rr < o.threshold) {
            o.done = 1;
            o.converged = 1;
            inner[inst].done = 1;
        } else if (!first && inner[inst].done && !inner[inst].converged) {
            // the last inner solve stopped at its iteration cap: further sweeps would repeat it
            o.done = 1;
            o.converged = 0;
            inner[inst].done = 1;
        }
    }
}

__global__ __launch_bounds__(kBatchThreads) void bacc_kernel(BatchTables t, const float* __restrict__ xf,
                                                             double* __restrict__ dv, const CgScalars* outer) {
    const int c = blockIdx.x, inst = t.chunk_inst[c];
    if (outer[inst].done) return;
    const int d0 = 3 * t.chunk_row0[c], d1 = 3 * t.chunk_row1[c];
    for (int i = d0 + (int)threadIdx.x; i < d1; i += blockDim.x) dv[i] += (double)xf[i];
}

template <typename TV, int G>
void iteration_g(const BatchCg& a, hipStream_t s) {
    const dim3 grid(a.t.grid > 0 ? a.t.grid : a.t.num_chunks), block(kBatchThreads);
    launch_k((bcg_apply_kernel<TV, G>), grid, block, 0, s, a.t, a.op, (const TV*)a.S, (const TV*)a.z, (TV*)a.p[0],
             (TV*)a.p[1], (TV*)a.q, a.partial, a.sc, (const int*)a.kA, a.kV);
    launch_k((bcg_vector_kernel<TV>), grid, block, 0, s, a.t, (TV*)a.x, (TV*)a.r, (TV*)a.z, (const TV*)a.p[0],
             (const TV*)a.p[1], (const TV*)a.q, (const TV*)a.dinv, a.partial, (const CgScalars*)a.sc,
             (const int*)a.kV, a.kA);
}

template <typename TV>
void iteration_t(const BatchCg& a, int lanes, hipStream_t s) {
    if (lanes == 4) iteration_g<TV, 4>(a, s);
    else if (lanes == 16) iteration_g<TV, 16>(a, s);
    else iteration_g<TV, 8>(a, s);
}

}  // namespace

void launch_batch_cg_init(const BatchCg& a, const double* b, const CgScalars* skip, double tol, int, hipStream_t s) {
    const dim3 grid(a.t.num_chunks), block(kBatchThreads);
    if (a.fp32)
        launch_k((bcg_init_kernel<float>), grid, block, 0, s, a.t, a.op, (const float*)a.S, b, (float*)a.x,
                 (float*)a.r, (float*)a.z, (float*)a.dinv, a.partial, skip);
    else
        launch_k((bcg_init_kernel<double>), grid, block, 0, s, a.t, a.op, (const double*)a.S, b, (double*)a.x,
                 (double*)a.r, (double*)a.z, (double*)a.dinv, a.partial, skip);
    launch_k(bcg_init2_kernel, dim3(a.t.num_instances), block, 0, s, a.t, (const double*)a.partial, a.sc, skip, tol,
             a.kA, a.kV);
}

void launch_batch_cg_iteration(const BatchCg& a, int lanes, hipStream_t s) {
    if (a.fp32) iteration_t<float>(a, lanes, s);
    else iteration_t<double>(a, lanes, s);
}

void launch_batch_residual(const BatchCg& a, const double* S64, const double* b, const double* dv, int first,
                           double* r64, CgScalars* outer, double tol, int lanes, hipStream_t s) {
    const dim3 grid(a.t.num_chunks), block(kBatchThreads);
    if (lanes == 4)
        launch_k(bres64_kernel<4>, grid, block, 0, s, a.t, a.op, S64, b, dv, first, r64, (const CgScalars*)outer,
                 a.partial);
    else if (lanes == 16)
        launch_k(bres64_kernel<16>, grid, block, 0, s, a.t, a.op, S64, b, dv, first, r64, (const CgScalars*)outer,
                 a.partial);
    else
        launch_k(bres64_kernel<8>, grid, block, 0, s, a.t, a.op, S64, b, dv, first, r64, (const CgScalars*)outer,
                 a.partial);
    launch_k(brefine_check_kernel, dim3(a.t.num_instances), block, 0, s, a.t, (const double*)a.partial, first, tol,
             outer, a.sc);
}

void launch_batch_accumulate(const BatchCg& a, const float* xf, double* dv, const CgScalars* outer, hipStream_t s) {
    launch_k(bacc_kernel, dim3(a.t.num_chunks), dim3(kBatchThreads), 0, s, a.t, xf, dv, outer);
}

}  // namespace femgpu
