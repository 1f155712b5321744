// Adjoint of one implicit-Euler step (one Newton iteration) on tetrahedra, for gradients of a loss
// with respect to the state and the per-element Young's moduli.
//
// Forward step (Math/OdeSolverEulerImplicit.cpp:55-161 with newState = current state):
//   S(x) dv = b(x, v, E),  b = f(x) - dt K(x) v + fg + fext (BC rows 0),
//   S = lower-triangle self-adjoint view of A = M/dt + dt K with BC rows/columns -> identity,
//   v' = v + dv,  x' = x + dt v'.
// Given the loss adjoints (lx, lv) of (x', v'):
//   gv  = lv + dt lx                         (adjoint of v', through x' = x + dt v')
//   S mu = gv,  mu~ = mu with BC entries 0   (S is symmetric: the transpose solve reuses S)
//   phi(x, v, E) = mu~ . [f(x) - dt K(x) v] - dt mu~ . S_K(x) dv     (mu~, dv held fixed)
//   lx_prev = lx + d phi / dx,  lv_prev = gv - dt K^T mu~,  dL/dE_e += phi_e / E_e
// (K and f are linear in E at fixed Poisson ratio, so d phi_e / d E_e = phi_e / E_e.)
// phi is a sum over elements; each element's d phi_e / d x_e (12 values) is computed in forward mode
// with dual numbers through the same factored corotational stiffness the forward kernel uses,
// including the rotation's derivative (polar_rotation on Dual).
#include "fem_adjoint.hpp"
#include "fem_device.hpp"

namespace femgpu {

namespace {

constexpr int kAdjLanes = 13;  // 12 tangent directions + 1 value pass per element

__device__ inline void load_element(const AdjointArgs& a, int e, int id[4], double X0[12], double xv[12],
                                    double vv[12], double mt[12], double dv[12]) {
    const int4 c = a.conn[e];
    id[0] = c.x;
    id[1] = c.y;
    id[2] = c.z;
    id[3] = c.w;
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t d = 3 * (size_t)id[n] + k;
            X0[3 * n + k] = a.x0[d];
            xv[3 * n + k] = a.x[d];
            vv[3 * n + k] = a.v[d];
            dv[3 * n + k] = a.v_next[d] - a.v[d];
            mt[3 * n + k] = a.bcmask[d] ? 0.0 : a.mu[d];
        }
}

// l^T K r for one 3x3 block
template <typename T>
__device__ __forceinline__ T bilinear(const double* l, const T* K, const double* r) {
    T s = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += (l[i] * r[j]) * K[3 * i + j];
    return s;
}

// l^T sym_lower(K) r, sym_lower(K)_ij = K_max(i,j),min(i,j) (the lower triangle mirrored)
template <typename T>
__device__ __forceinline__ T bilinear_lower(const double* l, const T* K, const double* r) {
    T s = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += (l[i] * r[j]) * (i >= j ? K[3 * i + j] : K[3 * j + i]);
    return s;
}

// Stiffness block K_pq of coro_block in dual arithmetic from the element-wide factors (R, W_p in duals; the
// gradients g and V, lambda, mu constants): K = V [lambda h_p h_q^T + mu h_q h_p^T + mu (g_p . g_q) I] plus
// W_p [g_q]x R^T when the correction applies; h = R g.
__device__ __forceinline__ void coro_block_dual(const Dual* R, const Dual* W, const Dual* hp, const Dual* hq,
                                                const double* gp, const double* gq, double V, double lam,
                                                double mu, bool correction, Dual* K) {
    const double gab = gp[0] * gq[0] + gp[1] * gq[1] + gp[2] * gq[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            K[3 * i + j] = V * (lam * hp[i] * hq[j] + mu * hq[i] * hp[j] + Dual((i == j) ? mu * gab : 0.0));
    if (correction) {
        Dual WG[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            WG[3 * i + 0] = W[3 * i + 1] * gq[2] - W[3 * i + 2] * gq[1];
            WG[3 * i + 1] = -W[3 * i + 0] * gq[2] + W[3 * i + 2] * gq[0];
            WG[3 * i + 2] = W[3 * i + 0] * gq[1] - W[3 * i + 1] * gq[0];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                K[3 * i + j] += WG[3 * i] * R[3 * j] + WG[3 * i + 1] * R[3 * j + 1] + WG[3 * i + 2] * R[3 * j + 2];
    }
}

// Row a's three values of a 12-vector (or of the 4 x 3 gradients) without indexing a register array by a
// loop variable (which would place the array in scratch memory).
__device__ __forceinline__ void row3(const double* v, int a, double* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = a == 0 ? v[i] : a == 1 ? v[3 + i] : a == 2 ? v[6 + i] : v[9 + i];
}

// One lane per (element, direction): directions 0..11 carry the tangent of x_e's coordinate dir and write
// d phi_e / d x_(e, dir); direction 12 (tangent zero) adds phi_e / E_e to dL/dE_e and writes -dt K_e^T mu~_e.
// The element-wide factors are formed once; the per-row factor W_p and the blocks K_pq follow row by row in
// loops kept rolled, and the node vectors (mu~, v, dv) are read where they are used, so the lane's live set
// stays in registers.
__global__ __launch_bounds__(256) void adj_coro_kernel(AdjointArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = t / kAdjLanes, dir = t % kAdjLanes;
    if (e >= a.num_elements) return;
    const int4 c4 = a.conn[e];
    double X0[12];
    Dual X[12];
    {
        const int id[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const size_t d = 3 * (size_t)id[n] + k;
                X0[3 * n + k] = a.x0[d];
                X[3 * n + k] = Dual(a.x[d], 3 * n + k == dir ? 1.0 : 0.0);
            }
    }
    double g[4][3];
    const double V = tet_rest_gradients(X0, g);
    const double* gf = &g[0][0];
    const double4 m = a.mat[e];
    const double lam = m.x, mu = m.y, dt = a.dt;
    CoroCommon<Dual> cc;
    coro_common<Dual>(X0, X, g, lam, mu, cc);
    auto node = [&](int p) { return p == 0 ? c4.x : p == 1 ? c4.y : p == 2 ? c4.z : c4.w; };
    auto mtilde = [&](int p, double* out) {  // mu~ of node p (zero on BC rows)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t d = 3 * (size_t)node(p) + k;
            out[k] = a.bcmask[d] ? 0.0 : a.mu[d];
        }
    };
    // mu~ . f(x)
    Dual phi(0.0);
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        double gp[3], mp[3];
        row3(gf, p, gp);
        mtilde(p, mp);
        Dual fp[3];
        coro_force<Dual>(cc, gp, V, fp);
        phi += mp[0] * fp[0] + mp[1] * fp[1] + mp[2] * fp[2];
    }
    double gv[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) gv[i] = 0.0;
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        double gp[3], mp[3], dvp[3];
        row3(gf, p, gp);
        mtilde(p, mp);
        const int ip = node(p);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t d = 3 * (size_t)ip + k;
            dvp[k] = a.v_next[d] - a.v[d];
        }
        Dual W[9], hp[3];
        coro_weight<Dual>(cc, gp, V, lam, mu, W);
        const Dual gpd[3] = {Dual(gp[0]), Dual(gp[1]), Dual(gp[2])};
        mv3(cc.R, gpd, hp);
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            double gq[3], mq[3], vq[3], dvq[3];
            row3(gf, q, gq);
            mtilde(q, mq);
            const int iq = node(q);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const size_t d = 3 * (size_t)iq + k;
                vq[k] = a.v[d];
                dvq[k] = a.v_next[d] - vq[k];
            }
            Dual hq[3], K[9];
            const Dual gqd[3] = {Dual(gq[0]), Dual(gq[1]), Dual(gq[2])};
            mv3(cc.R, gqd, hq);
            coro_block_dual(cc.R, W, hp, hq, gp, gq, V, lam, mu, cc.correction, K);
            phi -= dt * bilinear(mp, K, vq);
            if (p == q) {
                phi -= dt * bilinear_lower(mp, K, dvp);
            } else if (ip > iq) {
                // block (I, J), I > J, and its mirror (J, I) = K_pq^T in the self-adjoint view
                phi -= dt * (bilinear(mp, K, dvq) + bilinear(dvp, K, mq));
            }
            if (dir == 12) {
                double kq[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) kq[j] = -dt * (K[j].v * mp[0] + K[3 + j].v * mp[1] + K[6 + j].v * mp[2]);
                // gv[3 q + j] += kq[j], q selected without a dynamic register index
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
                    if (qq == q)
#pragma unroll
                        for (int j = 0; j < 3; ++j) gv[3 * qq + j] += kq[j];
            }
        }
    }
    if (dir < 12) {
        a.gx_rec[12 * (size_t)e + dir] = phi.d;
    } else {
        a.grad_E[e] += phi.v / m.w;
#pragma unroll
        for (int i = 0; i < 12; ++i) a.gv_rec[12 * (size_t)e + i] = gv[i];
    }
}

// Linear tets: K = K0 is constant and symmetric, f = -K0 (x - x0):
// d phi / dx = -K0 mu~, -dt K^T mu~ = -dt K0 mu~, phi = -mu~ . K0 ((x - x0) + dt v + dt dv).
__global__ __launch_bounds__(256) void adj_linear_kernel(AdjointArgs a) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.num_elements) return;
    int id[4];
    double X0[12], xv[12], vv[12], mt[12], dv[12];
    load_element(a, e, id, X0, xv, vv, mt, dv);
    double g[4][3];
    const double V = tet_rest_gradients(X0, g);
    const double4 m = a.mat[e];
    double w[12], Km[12];
    for (int i = 0; i < 12; ++i) {
        w[i] = (xv[i] - X0[i]) + a.dt * (vv[i] + dv[i]);
        Km[i] = 0.0;
    }
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) {
            double K[9];
            linear_block<double>(g, V, m.x, m.y, p, q, K);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Km[3 * p + i] += K[3 * i + j] * mt[3 * q + j];  // (K0 mu~)_p
        }
    double phi = 0.0;
    for (int i = 0; i < 12; ++i) phi -= Km[i] * w[i];
    for (int i = 0; i < 12; ++i) {
        a.gx_rec[12 * (size_t)e + i] = -Km[i];
        a.gv_rec[12 * (size_t)e + i] = -a.dt * Km[i];
    }
    a.grad_E[e] += phi / m.w;
}

// gv = lv + dt lx (the right-hand side of the adjoint solve)
__global__ __launch_bounds__(256) void adj_rhs_kernel(int n, const double* __restrict__ lx,
                                                      const double* __restrict__ lv, double dt,
                                                      double* __restrict__ gv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gv[i] = lv[i] + dt * lx[i];
}

// Per node, in element order: lx_out = lx + sum gx_rec, lv_out = gv + sum gv_rec.
__global__ __launch_bounds__(256) void adj_gather_kernel(int num_nodes, const int* __restrict__ node_ptr,
                                                         const int* __restrict__ node_list,
                                                         const double* __restrict__ gx_rec,
                                                         const double* __restrict__ gv_rec,
                                                         const double* lx, const double* __restrict__ gv,
                                                         double* lx_out, double* lv_out) {  // lx_out may alias lx
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= num_nodes) return;
    double sx[3] = {0.0, 0.0, 0.0}, sv[3] = {0.0, 0.0, 0.0};
    for (int k = node_ptr[n]; k < node_ptr[n + 1]; ++k) {
        const int code = node_list[k];
        const size_t base = 12 * (size_t)(code >> 2) + 3 * (code & 3);
        for (int c = 0; c < 3; ++c) {
            sx[c] += gx_rec[base + c];
            sv[c] += gv_rec[base + c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        const size_t d = 3 * (size_t)n + c;
        lx_out[d] = lx[d] + sx[c];
        lv_out[d] = gv[d] + sv[c];
    }
}

inline dim3 grid_for(int64_t n, int threads) { return dim3((unsigned)((n + threads - 1) / threads)); }

}  // namespace

void launch_adjoint_rhs(int n, const double* lx, const double* lv, double dt, double* gv, hipStream_t s) {
    hipLaunchKernelGGL(adj_rhs_kernel, grid_for(n, 256), dim3(256), 0, s, n, lx, lv, dt, gv);
}

void launch_adjoint_elements(int kind, const AdjointArgs& a, hipStream_t s) {
    if (kind == 1)
        hipLaunchKernelGGL(adj_coro_kernel, grid_for((int64_t)a.num_elements * kAdjLanes, 256), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(adj_linear_kernel, grid_for(a.num_elements, 256), dim3(256), 0, s, a);
}

void launch_adjoint_gather(int num_nodes, const int* node_ptr, const int* node_list, const double* gx_rec,
                           const double* gv_rec, const double* lx, const double* gv, double* lx_out, double* lv_out,
                           hipStream_t s) {
    hipLaunchKernelGGL(adj_gather_kernel, grid_for(num_nodes, 256), dim3(256), 0, s, num_nodes, node_ptr, node_list,
                       gx_rec, gv_rec, lx, gv, lx_out, lv_out);
}

}  // namespace femgpu
