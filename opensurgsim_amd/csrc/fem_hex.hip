// Hexahedral (Fem3DElementCube) path on gfx950: element-by-element operators over a greedy element
// colouring (no two elements of one colour share a node, so the scatter needs no atomics and the
// summation order is fixed).
//
//   FP32 inner CG:  cached A_e = M_e / dt + dt K_e as 36 packed 3x3 blocks (a <= b), one element per
//                   lane, stored interleaved by 64-element tiles so every load is wave-contiguous.
//   FP64 pieces:    matrix-free Gauss-point operator: (A p)_a = sum_g w_g detJ_g [rho N_a (sum_b N_b p_b) / dt
//                   + dt sigma(p) g_a], sigma = lambda tr(eps) I + 2 mu eps, eps = sym(sum_b p_b g_b^T);
//                   sigma g_a is the B_a^T C B p of Physics/Fem3DElementCube.cpp:213-288.
#include "fem_hex.hpp"
#include "fem_device.hpp"
#include "fem_reduce.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace femgpu {

namespace {

#define HCHECK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Node sign table of Physics/Fem3DElementCube.cpp:72-77 and the 2-point rule of
// Math/GaussLegendreQuadrature.cpp:30-34 (+1/sqrt(3) first; weights 1).
__host__ __device__ __forceinline__ double hex_sign(int i, int axis) {
    const int s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                         {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    return (double)s[i][axis];
}
__host__ __device__ __forceinline__ double gauss2(int i) {
    return i == 0 ? 0.57735026918962576451 : -0.57735026918962576451;
}

// Shape functions, their physical gradients and det J at Gauss point (gi, gj, gk).
// J(r, c) = dx_c / dxi_r; dN/dx = J^-1 dN/dxi (Physics/Fem3DElementCube.cpp:169-251).
__host__ __device__ __forceinline__ double hex_gp(const double* X, int gi, int gj, int gk, double N[8],
                                                  double g[8][3]) {
    const double xi[3] = {gauss2(gi), gauss2(gj), gauss2(gk)};
    double dN[8][3];
    for (int i = 0; i < 8; ++i) {
        const double f0 = 1 + xi[0] * hex_sign(i, 0), f1 = 1 + xi[1] * hex_sign(i, 1), f2 = 1 + xi[2] * hex_sign(i, 2);
        N[i] = 0.125 * f0 * f1 * f2;
        dN[i][0] = 0.125 * hex_sign(i, 0) * f1 * f2;
        dN[i][1] = 0.125 * f0 * hex_sign(i, 1) * f2;
        dN[i][2] = 0.125 * f0 * f1 * hex_sign(i, 2);
    }
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 8; ++i)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) J[3 * r + c] += X[3 * i + c] * dN[i][r];
    double Ji[9];
    const double d = inv3(J, Ji);
    for (int i = 0; i < 8; ++i)
        for (int r = 0; r < 3; ++r) g[i][r] = Ji[3 * r] * dN[i][0] + Ji[3 * r + 1] * dN[i][1] + Ji[3 * r + 2] * dN[i][2];
    return d;
}

// Block k (0..35) of the upper block triangle (a <= b), row-major over a.
__host__ __device__ constexpr int blk_a(int k) {
    int a = 0;
    while (k >= 8 - a) {
        k -= 8 - a;
        ++a;
    }
    return a;
}
__host__ __device__ constexpr int blk_b(int k) {
    int a = 0;
    while (k >= 8 - a) {
        k -= 8 - a;
        ++a;
    }
    return a + k;
}

// Cached FP32 element matrices, interleaved for one element per lane: position t of the padded colour
// order (tile T = t / 64, lane l = t % 64) keeps its 300 floats as 75 float4 chunks at A4[(75 T + k) 64 + l],
// so each wave-wide load reads 1 KiB contiguous. Floats 0..251: the 28 strictly upper node blocks (a < b,
// row-major over a, then b), 9 each, row-major; floats 252..299: the 8 diagonal blocks, the 6 entries
// (0,0) (0,1) (0,2) (1,1) (1,2) (2,2) of each (A_e is symmetric).
constexpr int kHexChunks = 75;
constexpr int kHexOffFloats = 252;
constexpr int kHexMaxDistinct = 32;  // distinct-matrix table: at most 38.4 KB of LDS per workgroup

// Strictly upper node block o (0..27) of the packed order: its nodes (a, b), a < b.
__host__ __device__ constexpr int off_a(int o) {
    int a = 0;
    while (o >= 7 - a) {
        o -= 7 - a;
        ++a;
    }
    return a;
}
__host__ __device__ constexpr int off_b(int o) {
    int a = 0;
    while (o >= 7 - a) {
        o -= 7 - a;
        ++a;
    }
    return a + 1 + o;
}
__host__ __device__ constexpr int off_index(int a, int b) { return 7 * a - a * (a - 1) / 2 + (b - a - 1); }
// Packed float of entry (i, j) of a diagonal block (upper triangle, i <= j).
__host__ __device__ constexpr int sym_index(int i, int j) { return i == 0 ? j : (i == 1 ? 2 + j : 5); }

// ------------------------------------------------------------------------------------------
// Set-up: A_e = cm M_e + ck K_e in FP32 (implicit Euler with Rayleigh damping: cm = 1/dt + alpha_M,
// ck = dt + alpha_K; lumped: M_e's row sums on the diagonal), one thread per (element, packed block).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void hex_prepare_kernel(int count_pad, const int* __restrict__ order_pad,
                                                          const int* __restrict__ conn,
                                                          const double* __restrict__ x0, const double4* __restrict__ mat,
                                                          double cm, double ck, int lumped, float* __restrict__ A32) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= (int64_t)count_pad * 36) return;
    const int t = (int)(tid / 36), kb = (int)(tid % 36);
    const int e = order_pad[t];
    if (e < 0) return;
    const int a = blk_a(kb), b = blk_b(kb);
    double X[24];
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) X[3 * i + c] = x0[3 * (size_t)conn[8 * e + i] + c];
    const double4 m = mat[e];
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
            for (int gk = 0; gk < 2; ++gk) {
                double N[8], g[8][3];
                const double w = hex_gp(X, gi, gj, gk, N, g);
                const double gab = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
                double nn = N[a] * N[b];
                if (lumped) nn = a == b ? N[a] * (((N[0] + N[1]) + (N[2] + N[3])) + ((N[4] + N[5]) + (N[6] + N[7]))) : 0.0;
                const double ms = w * m.z * nn * cm;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        A[3 * i + j] += ck * w * (m.x * g[a][i] * g[b][j] + m.y * g[b][i] * g[a][j] +
                                                  (i == j ? m.y * gab : 0.0)) +
                                        (i == j ? ms : 0.0);
            }
    float* base = A32 + ((size_t)(t >> 6) * kHexChunks * 64) * 4 + (size_t)(t & 63) * 4;
    auto put = [&](int f, double v) { base[(size_t)(f >> 2) * 256 + (f & 3)] = (float)v; };
    if (a < b) {
        for (int i = 0; i < 9; ++i) put(9 * off_index(a, b) + i, A[i]);
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) put(kHexOffFloats + 6 * a + sym_index(i, j), A[3 * i + j]);
    }
}

// FP64 element operator y = alpha_m M pm + alpha_k K p for one element (matrix-free; lumped: M's row
// sums on the diagonal).
__device__ __forceinline__ void hex_apply64(const double* X, const double4 m, const double* p, const double* pm,
                                            double am, double ak, int lumped, double* y) {
    for (int i = 0; i < 24; ++i) y[i] = 0.0;
    for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
            for (int gk = 0; gk < 2; ++gk) {
                double N[8], g[8][3];
                const double w = hex_gp(X, gi, gj, gk, N, g);
                double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                double np[3] = {0, 0, 0};
                for (int b = 0; b < 8; ++b)
                    for (int i = 0; i < 3; ++i) {
                        np[i] += N[b] * pm[3 * b + i];
                        for (int j = 0; j < 3; ++j) H[3 * i + j] += p[3 * b + i] * g[b][j];
                    }
                const double sn = ((N[0] + N[1]) + (N[2] + N[3])) + ((N[4] + N[5]) + (N[6] + N[7]));
                const double tr = H[0] + H[4] + H[8];
                double sig[9];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        sig[3 * i + j] = m.y * (H[3 * i + j] + H[3 * j + i]) + (i == j ? m.x * tr : 0.0);
                const double wk = ak * w, wm = am * w * m.z;
                for (int a = 0; a < 8; ++a)
                    for (int i = 0; i < 3; ++i)
                        y[3 * a + i] += wk * (sig[3 * i] * g[a][0] + sig[3 * i + 1] * g[a][1] + sig[3 * i + 2] * g[a][2]) +
                                        wm * N[a] * (lumped ? sn * pm[3 * a + i] : np[i]);
            }
}

// y[nodes] += sign * (alpha_m M + alpha_k K) w_e over the elements of one colour (w masked on BC DOFs
// when mask != null). With v != null, the right-hand side terms: K acts on w = (x - x0) + kv v and M on v.
__global__ __launch_bounds__(128) void hex_apply64_kernel(int count, const int* __restrict__ order,
                                                          const int* __restrict__ conn, const double* __restrict__ x0,
                                                          const double4* __restrict__ mat, const double* __restrict__ p,
                                                          const double* __restrict__ xcur, const double* __restrict__ v,
                                                          double kv, const uint8_t* __restrict__ mask, double am,
                                                          double ak, int lumped, double sign, double* __restrict__ y,
                                                          const CgScalars* __restrict__ sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || (sc && sc->done)) return;
    const int e = order[t];
    int ids[8];
    for (int i = 0; i < 8; ++i) ids[i] = conn[8 * e + i];
    double X[24], w[24], wm[24], out[24];
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) {
            const size_t d = 3 * (size_t)ids[i] + c;
            X[3 * i + c] = x0[d];
            if (v) {
                w[3 * i + c] = (xcur[d] - x0[d]) + kv * v[d];
                wm[3 * i + c] = v[d];
            } else {
                w[3 * i + c] = (mask && mask[d]) ? 0.0 : p[d];
                wm[3 * i + c] = w[3 * i + c];
            }
        }
    hex_apply64(X, mat[e], w, wm, am, ak, lumped, out);
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) y[3 * (size_t)ids[i] + c] += sign * out[3 * i + c];
}

// FP32 cached operator over one colour: q[nodes] += A_e p_e, one element per lane. p is zero on BC DOFs without
// a mask: hex_cg_init_kernel zeroes r there, and the BC rows of the operator are the identity (q = p on them),
// so r, z and p stay zero on those rows through the solve. The 36 blocks stream in 9 groups of 4 (36 floats =
// 9 float4 chunks per group). TABLE: A_e is entry table_idx[t] of the distinct-matrix table, staged in LDS by
// every workgroup (the same floats and the same arithmetic as the per-position cache).
template <bool TABLE>
__global__ __launch_bounds__(256) void hex_apply32_kernel(int count_pad, const int* __restrict__ order_pad,
                                                          const int* __restrict__ conn,
                                                          const float4* __restrict__ A4,
                                                          const float4* __restrict__ table,
                                                          const int* __restrict__ table_idx, int num_distinct,
                                                          const float* __restrict__ p,
                                                          const uint8_t* __restrict__ mask, float* __restrict__ q,
                                                          const CgScalars* __restrict__ sc) {
    extern __shared__ float4 lds_table[];
    if (sc && sc->done) return;  // uniform over the grid; no scalars when the operator is timed alone
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (TABLE) {
        for (int i = threadIdx.x; i < num_distinct * kHexChunks; i += blockDim.x) lds_table[i] = table[i];
        __syncthreads();
    }
    if (t >= count_pad) return;
    const int e = order_pad[t];
    if (e < 0) return;
    const int4* c4 = reinterpret_cast<const int4*>(conn) + 2 * (size_t)e;
    const int4 n0 = c4[0], n1 = c4[1];
    const int ids[8] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, n1.z, n1.w};
    float pv[24], y[24];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int d = 3 * ids[a] + c;
            pv[3 * a + c] = p[d];  // zero on BC DOFs (the solve starts from r = 0 there; BC rows are the identity)
            y[3 * a + c] = 0.f;
        }
    const float4* At = A4 + (size_t)(t >> 6) * kHexChunks * 64 + (t & 63);
    const float4* Tt = TABLE ? lds_table + table_idx[t] * kHexChunks : nullptr;
    // the strictly upper blocks, 4 per group of 9 chunks: y_a += B p_b and y_b += B^T p_a
#pragma unroll
    for (int g = 0; g < 7; ++g) {
        float w[36];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const float4 v = TABLE ? Tt[9 * g + k] : At[(size_t)(9 * g + k) * 64];
            w[4 * k] = v.x;
            w[4 * k + 1] = v.y;
            w[4 * k + 2] = v.z;
            w[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int a = off_a(4 * g + m), b = off_b(4 * g + m);
            const float* B = w + 9 * m;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                y[3 * a + i] += B[3 * i] * pv[3 * b] + B[3 * i + 1] * pv[3 * b + 1] + B[3 * i + 2] * pv[3 * b + 2];
#pragma unroll
            for (int j = 0; j < 3; ++j)
                y[3 * b + j] += B[j] * pv[3 * a] + B[3 + j] * pv[3 * a + 1] + B[6 + j] * pv[3 * a + 2];
        }
        // The next group's loads take their address from a value this group produced (a zero the compiler
        // cannot see through), so they are not hoisted above this group's products: without it the compiler
        // keeps most of the 75 chunks in flight at once (over 300 registers, 1 wave per SIMD) instead of about
        // two groups (210 registers, 2 waves per SIMD).
        {
            int zero;
            asm volatile("v_mov_b32 %0, 0" : "=v"(zero) : "v"(y[3 * off_b(4 * g + 3) + 2]));
            if (TABLE)
                Tt += zero;
            else
                At += zero;
        }
    }
    // the diagonal blocks, 4 per group of 6 chunks (symmetric, 6 floats each)
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        float w[24];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float4 v = TABLE ? Tt[kHexOffFloats / 4 + 6 * g + k] : At[(size_t)(kHexOffFloats / 4 + 6 * g + k) * 64];
            w[4 * k] = v.x;
            w[4 * k + 1] = v.y;
            w[4 * k + 2] = v.z;
            w[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int a = 4 * g + m;
            const float* B = w + 6 * m;
            const float p0 = pv[3 * a], p1 = pv[3 * a + 1], p2 = pv[3 * a + 2];
            y[3 * a] += B[0] * p0 + B[1] * p1 + B[2] * p2;
            y[3 * a + 1] += B[1] * p0 + B[3] * p1 + B[4] * p2;
            y[3 * a + 2] += B[2] * p0 + B[4] * p1 + B[5] * p2;
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) q[3 * ids[a] + c] += y[3 * a + c];
}

// 64-bit FNV-1a hash of each padded position's 300 cached floats (bit patterns, chunk order); 0 for padding.
__global__ __launch_bounds__(256) void hex_hash_kernel(int count_pad, const int* __restrict__ order_pad,
                                                       const float4* __restrict__ A4,
                                                       unsigned long long* __restrict__ hash) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count_pad) return;
    if (order_pad[t] < 0) {
        hash[t] = 0ull;
        return;
    }
    const float4* At = A4 + (size_t)(t >> 6) * kHexChunks * 64 + (t & 63);
    unsigned long long h = 1469598103934665603ull;
    for (int k = 0; k < kHexChunks; ++k) {
        const float4 v = At[(size_t)k * 64];
        const unsigned int w[4] = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                                   __float_as_uint(v.w)};
        for (int i = 0; i < 4; ++i) {
            h ^= w[i];
            h *= 1099511628211ull;
        }
    }
    hash[t] = h;
}

// bad[block] = 1 when a position of the block differs (bit patterns) from the representative of its table
// entry; bad is cleared by the caller.
__global__ __launch_bounds__(256) void hex_table_verify_kernel(int count_pad, const int* __restrict__ order_pad,
                                                               const float4* __restrict__ A4,
                                                               const int* __restrict__ table_idx,
                                                               const int* __restrict__ rep, int* __restrict__ bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int diff = 0;
    if (t < count_pad && order_pad[t] >= 0) {
        const int r = rep[table_idx[t]];
        const float4* At = A4 + (size_t)(t >> 6) * kHexChunks * 64 + (t & 63);
        const float4* Ar = A4 + (size_t)(r >> 6) * kHexChunks * 64 + (r & 63);
        for (int k = 0; k < kHexChunks; ++k) {
            const float4 a = At[(size_t)k * 64], b = Ar[(size_t)k * 64];
            diff |= (__float_as_uint(a.x) != __float_as_uint(b.x)) | (__float_as_uint(a.y) != __float_as_uint(b.y)) |
                    (__float_as_uint(a.z) != __float_as_uint(b.z)) | (__float_as_uint(a.w) != __float_as_uint(b.w));
        }
    }
    if (__syncthreads_or(diff) && threadIdx.x == 0) bad[blockIdx.x] = 1;
}

// table[u] = the cached floats of position rep[u], as kHexChunks consecutive float4 chunks.
__global__ __launch_bounds__(256) void hex_table_fill_kernel(int num_distinct, const int* __restrict__ rep,
                                                             const float4* __restrict__ A4,
                                                             float4* __restrict__ table) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_distinct * kHexChunks) return;
    const int u = i / kHexChunks, k = i - u * kHexChunks;
    const int r = rep[u];
    table[i] = A4[(size_t)(r >> 6) * kHexChunks * 64 + (size_t)k * 64 + (r & 63)];
}

// FP64 diagonal of A over one colour.
__global__ __launch_bounds__(128) void hex_diag_kernel(int count, const int* __restrict__ order,
                                                       const int* __restrict__ conn, const double* __restrict__ x0,
                                                       const double4* __restrict__ mat, double cm, double ck,
                                                       int lumped, double* __restrict__ diag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int e = order[t];
    int ids[8];
    double X[24], d[24];
    for (int i = 0; i < 8; ++i) {
        ids[i] = conn[8 * e + i];
        for (int c = 0; c < 3; ++c) X[3 * i + c] = x0[3 * (size_t)ids[i] + c];
    }
    for (int i = 0; i < 24; ++i) d[i] = 0.0;
    const double4 m = mat[e];
    for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
            for (int gk = 0; gk < 2; ++gk) {
                double N[8], g[8][3];
                const double w = hex_gp(X, gi, gj, gk, N, g);
                const double sn = ((N[0] + N[1]) + (N[2] + N[3])) + ((N[4] + N[5]) + (N[6] + N[7]));
                for (int a = 0; a < 8; ++a) {
                    const double gg = g[a][0] * g[a][0] + g[a][1] * g[a][1] + g[a][2] * g[a][2];
                    const double nn = lumped ? N[a] * sn : N[a] * N[a];
                    for (int i = 0; i < 3; ++i)
                        d[3 * a + i] += ck * w * (m.x * g[a][i] * g[a][i] + m.y * g[a][i] * g[a][i] + m.y * gg) +
                                        w * m.z * nn * cm;
                }
            }
    for (int a = 0; a < 8; ++a)
        for (int c = 0; c < 3; ++c) diag[3 * (size_t)ids[a] + c] += d[3 * a + c];
}

// b -= alpha diag(M) v per DOF (the Rayleigh mass force with the global lumped form's diagonal,
// Physics/FemRepresentation.cpp:583-587).
__global__ __launch_bounds__(256) void hex_diag_damp_kernel(int n, const double* __restrict__ d,
                                                            const double* __restrict__ v, double alpha,
                                                            double* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] -= alpha * (d[i] * v[i]);
}

__global__ __launch_bounds__(256) void hex_cast_diag_kernel(int n, const double* __restrict__ d64,
                                                            float* __restrict__ d32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d32[i] = (float)d64[i];
}

// q[row] += sum_k val[k] p[col[k]] over the localized external operator rows: one lane per row, every row listed
// once, launched after the colour passes on the same stream.
template <typename TV>
__global__ __launch_bounds__(256) void hex_ext_apply_kernel(int rows, const int* __restrict__ row,
                                                            const int* __restrict__ ptr, const int* __restrict__ col,
                                                            const TV* __restrict__ val, const TV* __restrict__ p,
                                                            TV* __restrict__ q, const CgScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    TV acc = TV(0);
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) acc += val[k] * p[col[k]];
    q[row[r]] += acc;
}

// The diagonal entries of those rows added to d (a copy of the element operator's diagonal).
__global__ __launch_bounds__(256) void hex_ext_diag_kernel(int rows, const int* __restrict__ row,
                                                           const int* __restrict__ ptr, const int* __restrict__ col,
                                                           const double* __restrict__ val, double* __restrict__ d) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int i = row[r];
    double acc = d[i];
    for (int k = ptr[r]; k < ptr[r + 1]; ++k)
        if (col[k] == i) acc += val[k];
    d[i] = acc;
}

// b = fg + fext (+ the step's localized force) before the element forces are subtracted; BC DOFs zeroed
// afterwards.
__global__ __launch_bounds__(256) void hex_rhs_init_kernel(int n, const double* __restrict__ fg,
                                                           const double* __restrict__ fext,
                                                           const double* __restrict__ fstep, double* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double v = fg[i] + (fext ? fext[i] : 0.0);
        if (fstep) v += fstep[i];
        b[i] = v;
    }
}
// Boundary rows of the right-hand side zeroed, the others divided by div (Runge-Kutta 4 solves M/dt a = f/dt).
__global__ __launch_bounds__(256) void hex_rhs_bc_kernel(int n, const uint8_t* __restrict__ mask, double div,
                                                         double* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = mask[i] ? 0.0 : (div == 1.0 ? b[i] : b[i] / div);
}

// ------------------------------------------------------------------------------------------
// CG pieces for the element-by-element operator (Eigen recurrence, as in fem_kernels.hip).
// ------------------------------------------------------------------------------------------
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void hex_cg_init_kernel(int n, const double* __restrict__ b,
                                                                 const TV* __restrict__ diag, TV* __restrict__ x,
                                                                 TV* __restrict__ r, TV* __restrict__ z,
                                                                 TV* __restrict__ dinv, const uint8_t* __restrict__ mask,
                                                                 double* __restrict__ partial,
                                                                 const CgScalars* __restrict__ skip) {
    __shared__ double lds[64];
    if (skip && skip->done) return;
    double bb = 0.0, rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const TV s = mask[i] ? TV(1) : diag[i];  // BC rows of the system matrix are the identity
        const TV d = (s != TV(0)) ? TV(1) / s : TV(1);
        const TV bi = mask[i] ? TV(0) : (TV)b[i];  // zero on BC rows: r, z, p and x stay zero there
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

// Iteration k, part 1: convergence test of k - 1, beta, p = z + beta p, q = 0.
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void hex_cg_pupdate_kernel(int n, int k, int max_iter,
                                                                    const TV* __restrict__ z, TV* __restrict__ p,
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
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        p[i] = (k == 0) ? z[i] : z[i] + bt * p[i];
        q[i] = TV(0);
    }
}

// BC rows of the operator are the identity: q_i = p_i there; partial p . q.
template <typename TV>
__global__ __launch_bounds__(kCgThreads) void hex_cg_bc_pq_kernel(int n, const uint8_t* __restrict__ mask,
                                                                  const TV* __restrict__ p, TV* __restrict__ q,
                                                                  double* __restrict__ partial,
                                                                  const CgScalars* __restrict__ sc) {
    __shared__ double lds[64];
    if (sc->done) return;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        TV qi = q[i];
        if (mask[i]) {
            qi = p[i];
            q[i] = qi;
        }
        acc += (double)p[i] * (double)qi;
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[2 * kCgBlocks + blockIdx.x] = s;
}

// FP64 residual r = b - S dv from q64 = A (masked dv); partial |r|^2 for the refinement test.
__global__ __launch_bounds__(kCgThreads) void hex_residual_kernel(int n, const uint8_t* __restrict__ mask,
                                                                  const double* __restrict__ b,
                                                                  const double* __restrict__ dv,
                                                                  const double* __restrict__ q64,
                                                                  double* __restrict__ r64,
                                                                  double* __restrict__ partial) {
    __shared__ double lds[64];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double e = b[i] - (mask[i] ? dv[i] : q64[i]);
        r64[i] = e;
        acc += e * e;
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[3 * kCgBlocks + blockIdx.x] = s;
}

__global__ __launch_bounds__(kCgThreads) void hex_refine_check_kernel(const double* __restrict__ partial, int first,
                                                                      double tol, CgScalars* outer, CgScalars* inner) {
    __shared__ double lds[64];
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

__global__ __launch_bounds__(kCgThreads) void hex_accumulate_kernel(int n, const float* __restrict__ xf,
                                                                    double* __restrict__ dv,
                                                                    const CgScalars* __restrict__ outer) {
    if (outer->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dv[i] += (double)xf[i];
}

__global__ void hex_eval_kernel(int first, int count, const int* conn, const double* x0all, const double* xall,
                                const double4* mat, double* Kout, double* fout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int e = first + t;
    double X[24], u[24];
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) {
            const size_t d = 3 * (size_t)conn[8 * e + i] + c;
            X[3 * i + c] = x0all[d];
            u[3 * i + c] = xall[d] - x0all[d];
        }
    const double4 m = mat[e];
    double f[24];
    for (int i = 0; i < 24; ++i) f[i] = 0.0;
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int gi = 0; gi < 2; ++gi)
                for (int gj = 0; gj < 2; ++gj)
                    for (int gk = 0; gk < 2; ++gk) {
                        double N[8], g[8][3];
                        const double w = hex_gp(X, gi, gj, gk, N, g);
                        const double gab = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                K[3 * i + j] += w * (m.x * g[a][i] * g[b][j] + m.y * g[b][i] * g[a][j] +
                                                     (i == j ? m.y * gab : 0.0));
                    }
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) {
                    if (Kout) Kout[(size_t)t * 576 + (3 * a + i) * 24 + 3 * b + j] = K[3 * i + j];
                    f[3 * a + i] -= K[3 * i + j] * u[3 * b + j];
                }
            }
        }
    if (fout)
        for (int i = 0; i < 24; ++i) fout[(size_t)t * 24 + i] = f[i];
}

inline dim3 blocks_for(int64_t n, int threads) { return dim3((unsigned)((n + threads - 1) / threads)); }

// Host-side: element data for validity and masses (Physics/Fem3DElementCube.cpp:321-366: V > 1e-12;
// evaluateJ asserts invertibility with the Eigen threshold 1e-12 on det J).
double host_hex_volume(const double* X, int64_t e) {
    double v = 0.0;
    for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
            for (int gk = 0; gk < 2; ++gk) {
                double N[8], g[8][3];
                const double d = hex_gp(X, gi, gj, gk, N, g);
                if (!(std::fabs(d) > 1e-12))
                    throw std::invalid_argument("Found a non invertible matrix J while computing Fem3DElementCube "
                                                "stiffness matrix (element " + std::to_string(e) + ")");
                v += d;
            }
    if (!(v > 1e-12))
        throw std::invalid_argument("Fem3DElementCube ill-defined, its volume is not positive (element " +
                                    std::to_string(e) + ")");
    return v;
}

}  // namespace

void hex_release(HexData& h) {
    for (void* p : {(void*)h.conn, (void*)h.order, (void*)h.order_pad, (void*)h.A32, (void*)h.diag64, (void*)h.diag32,
                    (void*)h.table, (void*)h.table_idx, (void*)h.mdiag_cons, (void*)h.mdiag_lump, (void*)h.diagx64,
                    (void*)h.diagx32})
        if (p) (void)hipFree(p);
    h.mdiag_cons = h.mdiag_lump = nullptr;
    h.diagx64 = nullptr;
    h.diagx32 = nullptr;
    h.conn = h.order = h.order_pad = nullptr;
    h.table = nullptr;
    h.table_idx = nullptr;
    h.num_distinct = 0;
    h.A32 = nullptr;
    h.diag64 = nullptr;
    h.diag32 = nullptr;
    h.ready = false;
}

void hex_setup(HexData& h, int64_t num_nodes, const double* x0, int64_t num_elements, const int32_t* conn,
               const double* mat, std::vector<double>& mass_per_node, std::vector<int32_t>& rowptr_b,
               std::vector<int32_t>& colidx_b, hipStream_t s) {
    hex_release(h);
    h.num_nodes = num_nodes;
    h.num_elements = num_elements;
    mass_per_node.assign(num_nodes, 0.0);
    for (int64_t e = 0; e < num_elements; ++e) {
        double X[24];
        for (int i = 0; i < 8; ++i)
            for (int c = 0; c < 3; ++c) X[3 * i + c] = x0[3 * (size_t)conn[8 * e + i] + c];
        const double vol = host_hex_volume(X, e);
        const double mass = vol * mat[4 * e + 2];  // FemElement::getMass (Physics/FemElement.cpp:106-109)
        for (int i = 0; i < 8; ++i) mass_per_node[conn[8 * e + i]] += mass / 8;
    }
    // node-pair block pattern (Physics/FemRepresentation.cpp:121-132)
    std::vector<int64_t> cnt(num_nodes + 1, 0);
    for (int64_t e = 0; e < num_elements; ++e)
        for (int a = 0; a < 8; ++a) cnt[conn[8 * e + a] + 1] += 8;
    for (int64_t i = 0; i < num_nodes; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> adj(cnt[num_nodes]);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t e = 0; e < num_elements; ++e)
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b) adj[pos[conn[8 * e + a]]++] = conn[8 * e + b];
    }
    rowptr_b.assign(num_nodes + 1, 0);
    colidx_b.clear();
    for (int64_t i = 0; i < num_nodes; ++i) {
        auto b = adj.begin() + cnt[i], en = adj.begin() + cnt[i + 1];
        std::sort(b, en);
        auto last = std::unique(b, en);
        colidx_b.insert(colidx_b.end(), b, last);
        if (colidx_b.size() > (size_t)INT32_MAX / 9) throw std::invalid_argument("pattern too large for int32 indices");
        rowptr_b[i + 1] = (int32_t)colidx_b.size();
    }
    // greedy element colouring: lowest colour unused by any element sharing a node
    std::vector<uint32_t> used(num_nodes, 0);
    std::vector<int> color(num_elements);
    int ncol = 0;
    for (int64_t e = 0; e < num_elements; ++e) {
        uint32_t m = 0;
        for (int a = 0; a < 8; ++a) m |= used[conn[8 * e + a]];
        int c = 0;
        while (c < 32 && (m >> c) & 1u) ++c;
        if (c == 32) throw std::invalid_argument("element colouring needs more than 32 colours");
        color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < 8; ++a) used[conn[8 * e + a]] |= 1u << c;
    }
    h.num_colors = ncol;
    h.color_ptr.assign(ncol + 1, 0);
    for (int64_t e = 0; e < num_elements; ++e) h.color_ptr[color[e] + 1]++;
    for (int c = 0; c < ncol; ++c) h.color_ptr[c + 1] += h.color_ptr[c];
    std::vector<int> order(num_elements);
    {
        std::vector<int> p(h.color_ptr.begin(), h.color_ptr.end() - 1);
        for (int64_t e = 0; e < num_elements; ++e) order[p[color[e]]++] = (int)e;
    }
    // padded order: each colour starts on a 64-element tile (-1 marks padding)
    h.pad_ptr.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; ++c) h.pad_ptr[c + 1] = h.pad_ptr[c] + (h.color_ptr[c + 1] - h.color_ptr[c] + 63) / 64 * 64;
    std::vector<int> order_pad(h.pad_ptr[ncol], -1);
    for (int c = 0; c < ncol; ++c)
        std::copy(order.begin() + h.color_ptr[c], order.begin() + h.color_ptr[c + 1], order_pad.begin() + h.pad_ptr[c]);
    HCHECK(hipMalloc(&h.conn, (size_t)num_elements * 8 * sizeof(int)));
    HCHECK(hipMalloc(&h.order, (size_t)num_elements * sizeof(int)));
    HCHECK(hipMalloc(&h.order_pad, order_pad.size() * sizeof(int)));
    HCHECK(hipMalloc(&h.A32, order_pad.size() * (size_t)kHexChunks * 4 * sizeof(float)));
    h.h_order_pad = order_pad;
    HCHECK(hipMemcpyAsync(h.order_pad, order_pad.data(), order_pad.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HCHECK(hipMalloc(&h.diag64, (size_t)num_nodes * 3 * sizeof(double)));
    HCHECK(hipMalloc(&h.diag32, (size_t)num_nodes * 3 * sizeof(float)));
    HCHECK(hipMemcpyAsync(h.conn, conn, (size_t)num_elements * 8 * sizeof(int), hipMemcpyHostToDevice, s));
    HCHECK(hipMemcpyAsync(h.order, order.data(), (size_t)num_elements * sizeof(int), hipMemcpyHostToDevice, s));
    HCHECK(hipStreamSynchronize(s));
    h.a32_lumped = -1;
    h.ready = true;
}

namespace {

// The distinct-matrix table of the cached operator (HexData::table), or none (num_distinct = 0) when the
// positions hold more than kHexMaxDistinct distinct matrices or FEM_HEX_TABLE=0 is set. Positions are
// grouped by a hash of their floats and every position is then compared with its group's representative
// on the device, so the table holds exactly the floats of the per-position cache.
void build_table(HexData& h, hipStream_t s) {
    h.num_distinct = 0;
    const char* env = getenv("FEM_HEX_TABLE");
    if (env && std::strcmp(env, "0") == 0) return;
    const int pad = h.pad_ptr.back();
    if (pad == 0) return;
    const float4* A4 = reinterpret_cast<const float4*>(h.A32);
    unsigned long long* d_hash = nullptr;
    HCHECK(hipMalloc(&d_hash, (size_t)pad * sizeof(unsigned long long)));
    std::vector<unsigned long long> hash(pad);
    hipLaunchKernelGGL(hex_hash_kernel, blocks_for(pad, 256), dim3(256), 0, s, pad, h.order_pad, A4, d_hash);
    const hipError_t e1 = hipMemcpyAsync(hash.data(), d_hash, (size_t)pad * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(d_hash);
    HCHECK(e1);
    HCHECK(e2);
    std::vector<int> idx(pad, 0), rep;
    {
        std::vector<std::pair<unsigned long long, int>> seen;  // (hash, entry), at most kHexMaxDistinct
        for (int t = 0; t < pad; ++t) {
            if (h.h_order_pad[t] < 0) continue;
            int u = -1;
            for (const auto& q : seen)
                if (q.first == hash[t]) {
                    u = q.second;
                    break;
                }
            if (u < 0) {
                if ((int)rep.size() == kHexMaxDistinct) return;
                u = (int)rep.size();
                seen.emplace_back(hash[t], u);
                rep.push_back(t);
            }
            idx[t] = u;
        }
    }
    if (rep.empty()) return;
    const int nd = (int)rep.size();
    if (!h.table) HCHECK(hipMalloc(&h.table, (size_t)kHexMaxDistinct * kHexChunks * sizeof(float4)));
    if (!h.table_idx) HCHECK(hipMalloc(&h.table_idx, (size_t)pad * sizeof(int)));
    const int nblk = (pad + 255) / 256;
    int *d_rep = nullptr, *d_bad = nullptr;
    HCHECK(hipMalloc(&d_rep, (size_t)nd * sizeof(int)));
    if (hipMalloc(&d_bad, (size_t)nblk * sizeof(int)) != hipSuccess) {
        (void)hipFree(d_rep);
        throw std::runtime_error("hipMalloc failed for the table check");
    }
    std::vector<int> bad(nblk, 0);
    hipError_t err = hipMemcpyAsync(h.table_idx, idx.data(), (size_t)pad * sizeof(int), hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipMemcpyAsync(d_rep, rep.data(), (size_t)nd * sizeof(int), hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipMemsetAsync(d_bad, 0, (size_t)nblk * sizeof(int), s);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(hex_table_verify_kernel, dim3(nblk), dim3(256), 0, s, pad, h.order_pad, A4, h.table_idx,
                           d_rep, d_bad);
        hipLaunchKernelGGL(hex_table_fill_kernel, blocks_for((int64_t)nd * kHexChunks, 256), dim3(256), 0, s, nd,
                           d_rep, A4, reinterpret_cast<float4*>(h.table));
        err = hipMemcpyAsync(bad.data(), d_bad, (size_t)nblk * sizeof(int), hipMemcpyDeviceToHost, s);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    (void)hipFree(d_rep);
    (void)hipFree(d_bad);
    HCHECK(err);
    for (int b : bad)
        if (b) return;  // a hash collision between different matrices: keep the per-position cache
    h.num_distinct = nd;
}

void hex_prepare(HexData& h, const double* x0, const double4* mat, double cm, double ck, int lumped, hipStream_t s) {
    const int pad = h.pad_ptr.back();
    hipLaunchKernelGGL(hex_prepare_kernel, blocks_for((int64_t)pad * 36, 256), dim3(256), 0, s, pad, h.order_pad,
                       h.conn, x0, mat, cm, ck, lumped, h.A32);
    const int n = (int)(3 * h.num_nodes);
    HCHECK(hipMemsetAsync(h.diag64, 0, (size_t)n * sizeof(double), s));
    for (int c = 0; c < h.num_colors; ++c) {
        const int cnt = h.color_ptr[c + 1] - h.color_ptr[c];
        hipLaunchKernelGGL(hex_diag_kernel, blocks_for(cnt, 128), dim3(128), 0, s, cnt, h.order + h.color_ptr[c],
                           h.conn, x0, mat, cm, ck, lumped, h.diag64);
    }
    hipLaunchKernelGGL(hex_cast_diag_kernel, blocks_for(n, 256), dim3(256), 0, s, n, h.diag64, h.diag32);
    build_table(h, s);
    h.a32_cm = cm;
    h.a32_ck = ck;
    h.a32_lumped = lumped;
}

// y += sign (am M + ak K) w over all colours (FP64, matrix-free).
void apply64(HexData& h, const double* x0, const double4* mat, const double* p, const double* xcur, const double* v,
             double kv, const uint8_t* mask, double am, double ak, int lumped, double sign, double* y, hipStream_t s) {
    for (int c = 0; c < h.num_colors; ++c) {
        const int cnt = h.color_ptr[c + 1] - h.color_ptr[c];
        hipLaunchKernelGGL(hex_apply64_kernel, blocks_for(cnt, 128), dim3(128), 0, s, cnt, h.order + h.color_ptr[c],
                           h.conn, x0, mat, p, xcur, v, kv, mask, am, ak, lumped, sign, y, (const CgScalars*)nullptr);
    }
}

// q += A_ext p with the step's localized external operator rows (none: nothing launched).
template <typename TV>
void ext_apply(const HexStepArgs& a, const TV* p, TV* q, const CgScalars* sc, hipStream_t s) {
    if (a.ext_rows <= 0) return;
    const TV* val;
    if constexpr (std::is_same<TV, float>::value)
        val = a.ext_val32;
    else
        val = a.ext_val;
    hipLaunchKernelGGL((hex_ext_apply_kernel<TV>), blocks_for(a.ext_rows, 256), dim3(256), 0, s, a.ext_rows, a.ext_row,
                       a.ext_ptr, a.ext_col, val, p, q, sc);
}

// Runs CG iterations with the element-by-element operator until the device reports done.
template <typename TV>
int run_hex_cg(HexData& h, const HexStepArgs& a, bool fp32, CgScalars* sc, CgScalars* h_sc, int& predicted,
               TV* x, TV* r, TV* z, TV* p, TV* q, TV* dinv, hipStream_t s) {
    const int n = (int)(3 * h.num_nodes);
    auto iteration = [&](int k) {
        hipLaunchKernelGGL((hex_cg_pupdate_kernel<TV>), dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, k, a.max_iter, z,
                           p, q, a.partial, sc);
        if (k >= a.max_iter) return;
        for (int c = 0; c < h.num_colors; ++c) {
            const int cnt = h.color_ptr[c + 1] - h.color_ptr[c];
            if (fp32) {
                const int c0 = h.pad_ptr[c], pc = h.pad_ptr[c + 1] - c0;
                const float4* A4 = reinterpret_cast<const float4*>(h.A32) + (size_t)c0 * kHexChunks;
                if (h.num_distinct > 0)
                    hipLaunchKernelGGL(hex_apply32_kernel<true>, blocks_for(pc, 256), dim3(256),
                                       (size_t)h.num_distinct * kHexChunks * sizeof(float4), s, pc, h.order_pad + c0,
                                       h.conn, A4, reinterpret_cast<const float4*>(h.table), h.table_idx + c0,
                                       h.num_distinct, (const float*)p, a.bcmask, (float*)q, (const CgScalars*)sc);
                else
                    hipLaunchKernelGGL(hex_apply32_kernel<false>, blocks_for(pc, 256), dim3(256), 0, s, pc,
                                       h.order_pad + c0, h.conn, A4, (const float4*)nullptr, (const int*)nullptr, 0,
                                       (const float*)p, a.bcmask, (float*)q, (const CgScalars*)sc);
            } else
                hipLaunchKernelGGL(hex_apply64_kernel, blocks_for(cnt, 128), dim3(128), 0, s, cnt,
                                   h.order + h.color_ptr[c], h.conn, a.x0, a.mat, (const double*)p, nullptr, nullptr,
                                   0.0, a.bcmask, h.a32_cm, h.a32_ck, h.a32_lumped, 1.0, (double*)q,
                                   (const CgScalars*)sc);
        }
        ext_apply<TV>(a, p, q, sc, s);
        hipLaunchKernelGGL((hex_cg_bc_pq_kernel<TV>), dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, a.bcmask, p, q,
                           a.partial, sc);
        launch_cg_vector_raw(fp32, n, k, x, r, z, p, q, dinv, a.partial, sc, s);
    };
    int k = 0;
    const int first = std::max(1, std::min(predicted + 1, a.max_iter + 1));
    for (; k < first; ++k) iteration(k);
    for (;;) {
        HCHECK(hipMemcpyAsync(h_sc, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
        HCHECK(hipStreamSynchronize(s));
        if (h_sc->done || k > a.max_iter) break;
        const int chunk = std::min(std::max(2, predicted / 8), a.max_iter + 1 - k);
        for (int j = 0; j < chunk; ++j, ++k) iteration(k);
    }
    predicted = h_sc->iters + 1;
    return h_sc->iters;
}

}  // namespace

double hex_time_operator(HexData& h, const HexStepArgs& a, bool fp32, int reps, hipStream_t s) {
    if (fp32 && (h.a32_lumped < 0 || !h.A32)) throw std::runtime_error("no FP32 operator cached yet: step first");
    void* p = a.vec[3];
    void* q = a.vec[4];
    auto apply = [&] {
        for (int c = 0; c < h.num_colors; ++c) {
            const int cnt = h.color_ptr[c + 1] - h.color_ptr[c];
            if (fp32) {
                const int c0 = h.pad_ptr[c], pc = h.pad_ptr[c + 1] - c0;
                const float4* A4 = reinterpret_cast<const float4*>(h.A32) + (size_t)c0 * kHexChunks;
                if (h.num_distinct > 0)
                    hipLaunchKernelGGL(hex_apply32_kernel<true>, blocks_for(pc, 256), dim3(256),
                                       (size_t)h.num_distinct * kHexChunks * sizeof(float4), s, pc, h.order_pad + c0,
                                       h.conn, A4, reinterpret_cast<const float4*>(h.table), h.table_idx + c0,
                                       h.num_distinct, (const float*)p, a.bcmask, (float*)q,
                                       (const CgScalars*)nullptr);
                else
                    hipLaunchKernelGGL(hex_apply32_kernel<false>, blocks_for(pc, 256), dim3(256), 0, s, pc,
                                       h.order_pad + c0, h.conn, A4, (const float4*)nullptr, (const int*)nullptr, 0,
                                       (const float*)p, a.bcmask, (float*)q, (const CgScalars*)nullptr);
            } else {
                hipLaunchKernelGGL(hex_apply64_kernel, blocks_for(cnt, 128), dim3(128), 0, s, cnt,
                                   h.order + h.color_ptr[c], h.conn, a.x0, a.mat, (const double*)p, nullptr, nullptr,
                                   0.0, a.bcmask, h.a32_cm, h.a32_ck, h.a32_lumped, 1.0, (double*)q,
                                   (const CgScalars*)nullptr);
            }
        }
    };
    apply();  // warm-up
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    double sum = 0.0;
    hipError_t err = hipSuccess;
    for (auto& e : ev)
        if (err == hipSuccess) err = hipEventCreate(&e);
    if (err == hipSuccess) {
        for (int r = 0; r < reps && err == hipSuccess; ++r) {
            err = hipEventRecord(ev[2 * r], s);
            apply();
            if (err == hipSuccess) err = hipEventRecord(ev[2 * r + 1], s);
        }
        if (err == hipSuccess) err = hipEventSynchronize(ev.back());
        for (int r = 0; r < reps && err == hipSuccess; ++r) {
            float t = 0.f;
            err = hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]);
            sum += t;
        }
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    HCHECK(err);
    return sum / reps;
}

void hex_step(HexData& h, const HexStepArgs& a, fem_stats& stats, hipStream_t s) {
    const fem_options& o = *a.opt;
    const double dt = o.dt;
    const int n = (int)(3 * h.num_nodes);
    const bool fp32 = o.precision == FEM_PRECISION_FP32_REFINED;
    // A = M (1/dt + alpha_M) + K (dt + alpha_K): implicit Euler with Rayleigh damping
    // (Math/OdeSolverEulerImplicit.cpp:148, Physics/FemRepresentation.cpp:510-518)
    const double cm = 1.0 / dt + o.rayleigh_mass, ck = dt + o.rayleigh_stiffness;
    const int lumped = o.mass_lumping ? 1 : 0;
    if (h.a32_cm != cm || h.a32_ck != ck || h.a32_lumped != lumped) hex_prepare(h, a.x0, a.mat, cm, ck, lumped, s);
    // rhs = f - dt K v - alpha_M M v - alpha_K K v + gravity + fext
    //     = -K ((x - x0) + (dt + alpha_K) v) - alpha_M M v + fg + fext, BC DOFs zeroed
    hipLaunchKernelGGL(hex_rhs_init_kernel, blocks_for(n, 256), dim3(256), 0, s, n, a.fg, a.fext, a.fstep, a.b);
    apply64(h, a.x0, a.mat, nullptr, a.x, a.v, dt + o.rayleigh_stiffness, nullptr, o.rayleigh_mass, 1.0, lumped, -1.0,
            a.b, s);
    hipLaunchKernelGGL(hex_rhs_bc_kernel, blocks_for(n, 256), dim3(256), 0, s, n, a.bcmask, 1.0, a.b);
    launch_ext_rhs(a.rhs_rows, a.rhs_row, a.rhs_ptr, a.rhs_col, a.rhs_k, a.rhs_f, a.rhs_dt, a.x, a.x, a.v, 0, 1.0,
                   a.b, s);  // - dt K_ext v
    hex_solve(h, a, stats, s);
    launch_update(n, a.dv, a.x, a.x, a.v, dt, a.invalid, s);
}

void hex_set_operator(HexData& h, const HexStepArgs& a, double cm, double ck, int lumped, hipStream_t s) {
    if (h.a32_cm != cm || h.a32_ck != ck || h.a32_lumped != lumped) hex_prepare(h, a.x0, a.mat, cm, ck, lumped, s);
}

const double* hex_mass_diagonal(HexData& h, const double* x0, const double4* mat, bool lumped, hipStream_t s) {
    double*& d = lumped ? h.mdiag_lump : h.mdiag_cons;
    if (d) return d;
    const int n = (int)(3 * h.num_nodes);
    HCHECK(hipMalloc(&d, (size_t)n * sizeof(double)));
    HCHECK(hipMemsetAsync(d, 0, (size_t)n * sizeof(double), s));
    for (int c = 0; c < h.num_colors; ++c) {
        const int cnt = h.color_ptr[c + 1] - h.color_ptr[c];
        hipLaunchKernelGGL(hex_diag_kernel, blocks_for(cnt, 128), dim3(128), 0, s, cnt, h.order + h.color_ptr[c],
                           h.conn, x0, mat, 1.0, 0.0, lumped ? 1 : 0, d);
    }
    return d;
}

void hex_rhs(HexData& h, const HexStepArgs& a, const double* x, const double* v, double kv, double am, int lumped,
             double div, hipStream_t s, const double* damp_d) {
    const int n = (int)(3 * h.num_nodes);
    hipLaunchKernelGGL(hex_rhs_init_kernel, blocks_for(n, 256), dim3(256), 0, s, n, a.fg, a.fext, a.fstep, a.b);
    if (damp_d) hipLaunchKernelGGL(hex_diag_damp_kernel, blocks_for(n, 256), dim3(256), 0, s, n, damp_d, v, am, a.b);
    apply64(h, a.x0, a.mat, nullptr, x, v, kv, nullptr, damp_d ? 0.0 : am, 1.0, lumped, -1.0, a.b, s);
    hipLaunchKernelGGL(hex_rhs_bc_kernel, blocks_for(n, 256), dim3(256), 0, s, n, a.bcmask, div, a.b);
}

void hex_solve(HexData& h, const HexStepArgs& a, fem_stats& stats, hipStream_t s) {
    const fem_options& o = *a.opt;
    const int n = (int)(3 * h.num_nodes);
    const bool fp32 = o.precision == FEM_PRECISION_FP32_REFINED;
    const double cm = h.a32_cm, ck = h.a32_ck;
    const int lumped = h.a32_lumped;
    CgScalars* sc = a.sc;
    const double* diag64 = h.diag64;
    const float* diag32 = h.diag32;
    if (a.ext_rows > 0) {  // the Jacobi diagonal of the operator with the localized external entries
        if (!h.diagx64) {
            HCHECK(hipMalloc(&h.diagx64, (size_t)n * sizeof(double)));
            HCHECK(hipMalloc(&h.diagx32, (size_t)n * sizeof(float)));
        }
        HCHECK(hipMemcpyAsync(h.diagx64, h.diag64, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(hex_ext_diag_kernel, blocks_for(a.ext_rows, 256), dim3(256), 0, s, a.ext_rows, a.ext_row,
                           a.ext_ptr, a.ext_col, a.ext_val, h.diagx64);
        hipLaunchKernelGGL(hex_cast_diag_kernel, blocks_for(n, 256), dim3(256), 0, s, n, h.diagx64, h.diagx32);
        diag64 = h.diagx64;
        diag32 = h.diagx32;
    }
    if (!fp32) {
        double* x = a.dv;
        hipLaunchKernelGGL((hex_cg_init_kernel<double>), dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, a.b, diag64,
                           x, (double*)a.vec[1], (double*)a.vec[2], (double*)a.vec[5], a.bcmask, a.partial,
                           (const CgScalars*)nullptr);
        launch_cg_init2(a.partial, sc, nullptr, o.cg_tolerance, s);
        const int it = run_hex_cg<double>(h, a, false, sc, a.h_sc, a.predicted[0], x, (double*)a.vec[1],
                                          (double*)a.vec[2], (double*)a.vec[3], (double*)a.vec[4],
                                          (double*)a.vec[5], s);
        stats.last_cg_iterations = it;
        stats.total_cg_iterations += it;
        stats.last_cg_error = a.h_sc->err;
        stats.converged = a.h_sc->converged;
        stats.last_refinements = 0;
    } else {
        CgScalars* outer = sc + 1;
        HCHECK(hipMemsetAsync(a.dv, 0, (size_t)n * sizeof(double), s));
        int total = 0, sweeps = 0;
        for (int sweep = 0; sweep <= o.max_refinements; ++sweep) {
            if (sweep == 0) {
                hipLaunchKernelGGL(hex_residual_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, a.bcmask, a.b,
                                   a.dv, a.dv, a.r64, a.partial);  // dv = 0: r = b
            } else {
                double* q64 = (double*)a.vec[6];
                HCHECK(hipMemsetAsync(q64, 0, (size_t)n * sizeof(double), s));
                apply64(h, a.x0, a.mat, a.dv, nullptr, nullptr, 0.0, a.bcmask, cm, ck, lumped, 1.0, q64, s);
                ext_apply<double>(a, a.dv, q64, nullptr, s);
                hipLaunchKernelGGL(hex_residual_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, a.bcmask, a.b,
                                   a.dv, q64, a.r64, a.partial);
            }
            hipLaunchKernelGGL(hex_refine_check_kernel, dim3(1), dim3(kCgThreads), 0, s, a.partial, sweep == 0,
                               o.cg_tolerance, outer, sc);
            if (sweep == o.max_refinements) break;
            hipLaunchKernelGGL((hex_cg_init_kernel<float>), dim3(kCgBlocks), dim3(kCgThreads), 0, s, n, a.r64,
                               diag32, (float*)a.vec[0], (float*)a.vec[1], (float*)a.vec[2], (float*)a.vec[5],
                               a.bcmask, a.partial, (const CgScalars*)outer);
            launch_cg_init2(a.partial, sc, outer, o.inner_tolerance, s, kInnerFloor * o.cg_tolerance);
            const int ps = 1 + std::min(sweep, 15);
            const int it = run_hex_cg<float>(h, a, true, sc, a.h_sc, a.predicted[ps], (float*)a.vec[0],
                                             (float*)a.vec[1], (float*)a.vec[2], (float*)a.vec[3], (float*)a.vec[4],
                                             (float*)a.vec[5], s);
            HCHECK(hipMemcpyAsync(a.h_sc + 1, outer, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
            HCHECK(hipStreamSynchronize(s));
            if (a.h_sc[1].done) break;
            total += it + 1;
            hipLaunchKernelGGL(hex_accumulate_kernel, dim3(kCgBlocks), dim3(kCgThreads), 0, s, n,
                               (const float*)a.vec[0], a.dv, outer);
            sweeps++;
        }
        CgScalars oh;
        HCHECK(hipMemcpyAsync(&oh, outer, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
        HCHECK(hipStreamSynchronize(s));
        stats.last_cg_iterations = total;
        stats.total_cg_iterations += total;
        stats.last_cg_error = oh.err;
        stats.converged = oh.converged;
        stats.last_refinements = sweeps;
    }
}

void hex_eval_elements(HexData& h, int first, int count, const double* x0, const double* x, const double4* mat,
                       double* R, double* K, double* f, hipStream_t s) {
    if (count <= 0) return;
    if (R) {
        // rotation of a linear element: identity
        std::vector<double> I(9 * (size_t)count, 0.0);
        for (int t = 0; t < count; ++t) I[9 * t] = I[9 * t + 4] = I[9 * t + 8] = 1.0;
        HCHECK(hipMemcpyAsync(R, I.data(), I.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HCHECK(hipStreamSynchronize(s));
    }
    hipLaunchKernelGGL(hex_eval_kernel, blocks_for(count, 64), dim3(64), 0, s, first, count, h.conn, x0, x, mat, K, f);
}

}  // namespace femgpu
