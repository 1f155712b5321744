// Device-side small dense algebra and the Fem3D element kernels' per-element math (gfx950).
//
// Element algebra restated from the reference (paths relative to the reference root):
//   linear tet        Physics/Fem3DElementTetrahedron.cpp:92-235
//   corotational tet  Physics/Fem3DElementCorotationalTetrahedron.cpp:48-288
// in a factored form that never builds the reference's dense 12x12 intermediates:
//   g_a       = rest shape-function gradients (rows of V^-1 without the homogeneous column)
//   K0_ab     = V [lambda g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I]          (= V B^T E B block)
//   F         = sum_a x_a g_a^T = R S                                          (polar decomposition)
//   omega_(b,j) = (tr(S) I - S)^-1 (g_b x R^T e_j)                             (dR/dx_(b,j) = R [omega]x)
//   K_ab      = R K0_ab R^T + W_a [g_b]x R^T,
//   W_a       = R (-[q_a]x + V mu ([S g_a]x - S [g_a]x)) (tr(S) I - S)^-1,
//   q_a       = R (K0 (R^T (R^T x - x0)))_a  -- the reference evaluates the first Barbic term with
//               m_K = R K0 R^T already written (:203 then :276), which this reproduces,
//   f_a       = -V R [lambda tr(S - I) I + 2 mu (S - I)] g_a                   (= -m_K (x - R x0))
// Each identity is exact in real arithmetic; the oracle evaluates the reference's literal form and
// the tests compare both per element.
#pragma once

#include <hip/hip_runtime.h>

namespace femgpu {

// Forward-mode dual number (value, one tangent) for the adjoint's element derivatives.
struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0.0), d(0.0) {}
    __host__ __device__ explicit Dual(double value, double tangent = 0.0) : v(value), d(tangent) {}
    __host__ __device__ Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
    __host__ __device__ Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
    __host__ __device__ Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
    const double iv = 1.0 / b.v;
    return Dual(a.v * iv, (a.d - a.v * iv * b.d) * iv);
}
__host__ __device__ inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
__host__ __device__ inline bool operator<(Dual a, double b) { return a.v < b; }
__host__ __device__ inline bool operator>(Dual a, double b) { return a.v > b; }
__host__ __device__ inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
using ::fabs;
__host__ __device__ inline double value_of(double a) { return a; }
__host__ __device__ inline double value_of(Dual a) { return a.v; }

template <typename T>
__host__ __device__ __forceinline__ void cross3(const T* a, const T* b, T* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__host__ __device__ __forceinline__ T dot3(const T* a, const T* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// C = A B (3x3 row-major)
template <typename T>
__host__ __device__ __forceinline__ void mm3(const T* A, const T* B, T* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
// C = A^T B
template <typename T>
__host__ __device__ __forceinline__ void mtm3(const T* A, const T* B, T* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
// y = A x
template <typename T>
__host__ __device__ __forceinline__ void mv3(const T* A, const T* x, T* y) {
    y[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
    y[1] = A[3] * x[0] + A[4] * x[1] + A[5] * x[2];
    y[2] = A[6] * x[0] + A[7] * x[1] + A[8] * x[2];
}
template <typename T>
__host__ __device__ __forceinline__ T det3(const T* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}
// Cofactor inverse; returns det. inv untouched when det == 0.
template <typename T>
__host__ __device__ __forceinline__ T inv3(const T* A, T* inv) {
    T c00 = A[4] * A[8] - A[5] * A[7];
    T c01 = A[5] * A[6] - A[3] * A[8];
    T c02 = A[3] * A[7] - A[4] * A[6];
    T d = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (d == T(0)) return d;
    T id = T(1) / d;
    inv[0] = c00 * id;
    inv[3] = c01 * id;
    inv[6] = c02 * id;
    inv[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    inv[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    inv[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    inv[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    inv[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    inv[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return d;
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Golub & Van Loan, Alg. 8.4.3).
// A (symmetric, row-major) is overwritten; on exit ev[k] are eigenvalues, V columns eigenvectors.
__device__ inline void jacobi_eig3(double* A, double* ev, double* V) {
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        double diag = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int pq = 0; pq < 3; ++pq) {
            int p = pq == 2 ? 1 : 0;
            int q = pq == 0 ? 1 : 2;
            double apq = A[3 * p + q];
            if (apq == 0.0) continue;
            double app = A[3 * p + p], aqq = A[3 * q + q];
            double theta = (aqq - app) / (2.0 * apq);
            double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            double c = 1.0 / sqrt(t * t + 1.0);
            double s = t * c;
            // A <- J^T A J with J the rotation in the (p, q) plane
            for (int k = 0; k < 3; ++k) {
                double akp = A[3 * k + p], akq = A[3 * k + q];
                A[3 * k + p] = c * akp - s * akq;
                A[3 * k + q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = A[3 * p + k], aqk = A[3 * q + k];
                A[3 * p + k] = c * apk - s * aqk;
                A[3 * q + k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                double vkp = V[3 * k + p], vkq = V[3 * k + q];
                V[3 * k + p] = c * vkp - s * vkq;
                V[3 * k + q] = s * vkp + c * vkq;
            }
        }
    }
    ev[0] = A[0];
    ev[1] = A[4];
    ev[2] = A[8];
}

// Polar decomposition F = R S with Eigen's Transform::computeRotationScaling convention: R a proper
// rotation; when det F < 0 the sign flip goes to the largest singular value (S then has one negative
// eigenvalue). For det F > 0 this is the orthogonal polar factor. Near a rotation (|F^T F - I|_F < 1/2)
// it comes from the inverse-free Newton-Schulz iteration X <- X (3 I - X^T X) / 2 (Higham 2008, Functions
// of Matrices, 8.3), which converges quadratically there; otherwise from the scaled Newton iteration
// X <- (g X + X^-T / g) / 2 (Higham 1986, Frobenius scaling); and for det F <= 0 or a degenerate F
// through the eigen-decomposition of F^T F. Returns 1 if the eigen path was taken.
__device__ inline int polar_rotation(const double* F, double* R) {
    double d = det3(F);
    if (d > 0.0) {
        double C[9];
        mtm3(F, F, C);
        double e0 = 0.0;
        for (int i = 0; i < 9; ++i) {
            const double e = C[i] - ((i % 4 == 0) ? 1.0 : 0.0);
            e0 += e * e;
        }
        if (e0 < 0.25) {
            double X[9];
            for (int i = 0; i < 9; ++i) X[i] = F[i];
            for (int it = 0; it < 16; ++it) {
                double E[9];
                mtm3(X, X, E);
                double en = 0.0;
                for (int i = 0; i < 9; ++i) {
                    E[i] -= (i % 4 == 0) ? 1.0 : 0.0;
                    en += E[i] * E[i];
                }
                double Y[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        Y[3 * i + j] = X[3 * i + j] - 0.5 * (X[3 * i] * E[j] + X[3 * i + 1] * E[3 + j] +
                                                             X[3 * i + 2] * E[6 + j]);
                for (int i = 0; i < 9; ++i) X[i] = Y[i];
                if (en < 1e-20) break;  // this step took |X^T X - I| from below 1e-10 to the rounding level
            }
            for (int i = 0; i < 9; ++i) R[i] = X[i];
            return 0;
        }
    }
    double fn2 = 0.0;
    for (int i = 0; i < 9; ++i) fn2 += F[i] * F[i];
    if (d > 1e-6 * fn2 * sqrt(fn2)) {
        double X[9];
        for (int i = 0; i < 9; ++i) X[i] = F[i];
        bool scale = true;
        for (int it = 0; it < 40; ++it) {
            double Xi[9];
            double dx = inv3(X, Xi);
            (void)dx;
            double g = 1.0;
            if (scale) {
                double n1 = 0.0, n2 = 0.0;
                for (int i = 0; i < 9; ++i) {
                    n1 += X[i] * X[i];
                    n2 += Xi[i] * Xi[i];
                }
                g = sqrt(sqrt(n2 / n1));
            }
            double delta = 0.0;
            double Y[9];
            const double ig = 1.0 / g;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    // X^-T (i, j) = Xi (j, i)
                    double y = 0.5 * (g * X[3 * i + j] + Xi[3 * j + i] * ig);
                    double e = y - X[3 * i + j];
                    delta += e * e;
                    Y[3 * i + j] = y;
                }
            for (int i = 0; i < 9; ++i) X[i] = Y[i];
            if (delta < 1e-4) scale = false;
            if (delta < 1e-30) break;
        }
        for (int i = 0; i < 9; ++i) R[i] = X[i];
        return 0;
    }
    // Inverted or degenerate element: SVD through the eigen-decomposition of F^T F.
    double C[9], ev[3], V[9];
    mtm3(F, F, C);
    jacobi_eig3(C, ev, V);
    int o[3] = {0, 1, 2};
    if (ev[o[0]] < ev[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    if (ev[o[1]] < ev[o[2]]) { int t = o[1]; o[1] = o[2]; o[2] = t; }
    if (ev[o[0]] < ev[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    double v[3][3], u[3][3], sg[3];
    for (int k = 0; k < 3; ++k) {
        for (int i = 0; i < 3; ++i) v[k][i] = V[3 * i + o[k]];
        sg[k] = sqrt(fmax(ev[o[k]], 0.0));
    }
    // Right-handed V
    double c01[3];
    cross3(v[0], v[1], c01);
    if (dot3(c01, v[2]) < 0) for (int i = 0; i < 3; ++i) v[2][i] = -v[2][i];
    for (int k = 0; k < 2; ++k) {
        double fv[3];
        mv3(F, v[k], fv);
        double n = sqrt(dot3(fv, fv));
        if (n > 0) for (int i = 0; i < 3; ++i) u[k][i] = fv[i] / n;
        else for (int i = 0; i < 3; ++i) u[k][i] = (i == k) ? 1.0 : 0.0;
    }
    // Orthonormalise u1 against u0, complete with u2 = u0 x u1 (det U = +1).
    double p = dot3(u[0], u[1]);
    for (int i = 0; i < 3; ++i) u[1][i] -= p * u[0][i];
    double n1 = sqrt(dot3(u[1], u[1]));
    if (n1 > 0) for (int i = 0; i < 3; ++i) u[1][i] /= n1;
    cross3(u[0], u[1], u[2]);
    // F v2 = s2 u2 with s2 signed; det F < 0 <=> s2 < 0 here (det U = det V = 1).
    double fv2[3];
    mv3(F, v[2], fv2);
    double s2 = dot3(fv2, u[2]);
    double sgn0 = 1.0, sgn2 = 1.0;
    if (s2 < 0) {
        // Eigen: U_E = [u0, u1, -u2]; R = U_E diag(-1, 1, 1) V^T = [-u0, u1, -u2] V^T
        sgn0 = -1.0;
        sgn2 = -1.0;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            R[3 * i + j] = sgn0 * u[0][i] * v[0][j] + u[1][i] * v[1][j] + sgn2 * u[2][i] * v[2][j];
    (void)sg;
    return 1;
}

// Polar rotation of a dual-number F: the value as above; the tangent from differentiating F = R S
// with S symmetric. dR = R [w]x and R^T dF = [w]x S + dS give R^T dF - dF^T R = [w]x S + S [w]x
// = [(tr(S) I - S) w]x, so w = (tr(S) I - S)^-1 axial(R^T dF - dF^T R). Exact wherever tr(S) I - S is
// invertible (it is singular only when two principal stretches sum to zero); the tangent is set to zero
// otherwise, where the element's stiffness correction is dropped as well.
__device__ inline int polar_rotation(const Dual* F, Dual* R) {
    double Fv[9], Rv[9], dF[9];
    for (int i = 0; i < 9; ++i) {
        Fv[i] = F[i].v;
        dF[i] = F[i].d;
    }
    const int path = polar_rotation(Fv, Rv);
    double RtF[9], S[9], G[9], Gi[9], M[9];
    mtm3(Rv, Fv, RtF);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[3 * i + j] = 0.5 * (RtF[3 * i + j] + RtF[3 * j + i]);
    const double trS = S[0] + S[4] + S[8];
    for (int i = 0; i < 9; ++i) G[i] = ((i % 4 == 0) ? trS : 0.0) - S[i];
    double w[3] = {0.0, 0.0, 0.0};
    if (fabs(inv3(G, Gi)) > 1e-300) {
        mtm3(Rv, dF, M);  // M = R^T dF; A = M - M^T, axial(A) = (A_21, A_02, A_10)
        const double a[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
        mv3(Gi, a, w);
    }
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double dR[9];
    mm3(Rv, W, dR);
    for (int i = 0; i < 9; ++i) R[i] = Dual(Rv[i], dR[i]);
    return path;
}

// Per-element geometry of a tetrahedron at rest: gradients g[4][3] and signed volume V
// (Physics/Fem3DElementTetrahedron.cpp:218-359; g_a are the first three entries of row a of V^-1,
// Physics/Fem3DElementCorotationalTetrahedron.cpp:56-72, 254).
__device__ inline double tet_rest_gradients(const double* x0, double g[4][3]) {
    double e1[3], e2[3], e3[3];
    for (int c = 0; c < 3; ++c) {
        e1[c] = x0[3 + c] - x0[c];
        e2[c] = x0[6 + c] - x0[c];
        e3[c] = x0[9 + c] - x0[c];
    }
    double c23[3], c31[3], c12[3];
    cross3(e2, e3, c23);
    cross3(e3, e1, c31);
    cross3(e1, e2, c12);
    double det = dot3(e1, c23);
    double id = 1.0 / det;
    for (int c = 0; c < 3; ++c) {
        g[1][c] = c23[c] * id;
        g[2][c] = c31[c] * id;
        g[3][c] = c12[c] * id;
        g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
    }
    return det / 6.0;
}

// Flags reported per element.
enum : int { ELEM_FLAG_DET_R = 1, ELEM_FLAG_G_SINGULAR = 2, ELEM_FLAG_POLAR_EIG = 4 };

// Inverse of a node's 3x3 diagonal block of the FP32 operator (block db, row-major), formed in FP64 and
// rounded to float: the block-diagonal (block-Jacobi) preconditioner of the FP32 inner solves. A missing
// block (db < 0, a node in no element) or a numerically singular one falls back to the rule of Eigen's
// DiagonalPreconditioner, 1 / a_ii (1 where a_ii = 0).
__device__ inline void block_jacobi_inverse(const float* __restrict__ S, int db, float* Bf) {
    double A[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (db >= 0)
        for (int k = 0; k < 9; ++k) A[k] = (double)S[9 * (int64_t)db + k];
    double Ai[9];
    const double det = inv3(A, Ai);
    const double dscale = fabs(A[0] * A[4] * A[8]);
    if (!(fabs(det) > 1e-12 * dscale) || !(dscale > 0.0)) {
        for (int k = 0; k < 9; ++k) Ai[k] = 0.0;
        for (int c = 0; c < 3; ++c) Ai[4 * c] = A[4 * c] != 0.0 ? 1.0 / A[4 * c] : 1.0;
    }
    for (int k = 0; k < 9; ++k) Bf[k] = (float)Ai[k];
}

// Shared per-element state of the corotational stiffness, built once and consumed block by block.
template <typename TK>
struct CoroFactors {
    TK R[9];      // rotation
    TK W[4][9];   // W_a
    TK h[4][3];   // R g_a
    TK g[4][3];   // g_a
    TK V, lam, mu;
    bool correction;  // false when G is singular (or det R invalid): K = R K0 R^T only
};

// The element-wide part of the corotational factors: F, R, S, E = S - I, R sigma, (tr S I - S)^-1 and the
// element flags. T is the arithmetic (double; Dual for the adjoint's derivative with respect to x).
template <typename T>
struct CoroCommon {
    T R[9], S[9], Es[9], Rsig[9], G0i[9];
    bool correction;
    int flags;
};

template <typename T>
__device__ inline void coro_common(const double* x0, const T* x, const double g[4][3], double lam, double mu,
                                   CoroCommon<T>& c) {
    c.flags = 0;
    // F = I + sum_b (x_b - x0_b) g_b^T  (sum_b x0_b g_b^T = I; the displacement form keeps FP64 accuracy)
    T F[9] = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)};
    for (int b = 0; b < 4; ++b) {
        T u[3] = {x[3 * b] - x0[3 * b], x[3 * b + 1] - x0[3 * b + 1], x[3 * b + 2] - x0[3 * b + 2]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) F[3 * i + j] += u[i] * g[b][j];
    }
    T* R = c.R;
    if (polar_rotation(F, R)) c.flags |= ELEM_FLAG_POLAR_EIG;
    T detR = det3(R);
    // S = sym(R^T F);  E = S - I computed as sym(R^T (F - I) + (R^T - I))
    T RtF[9];
    mtm3(R, F, RtF);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c.S[3 * i + j] = 0.5 * (RtF[3 * i + j] + RtF[3 * j + i]);
    T Fm[9];
    for (int i = 0; i < 9; ++i) Fm[i] = F[i] - ((i % 4 == 0) ? 1.0 : 0.0);
    T RtD[9];
    mtm3(R, Fm, RtD);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T rt_ij = R[3 * j + i] - (i == j ? 1.0 : 0.0);
            T rt_ji = R[3 * i + j] - (i == j ? 1.0 : 0.0);
            c.Es[3 * i + j] = 0.5 * (RtD[3 * i + j] + RtD[3 * j + i] + rt_ij + rt_ji);
        }
    T trE = c.Es[0] + c.Es[4] + c.Es[8];
    // sigma = lambda tr(E) I + 2 mu E
    T sig[9];
    for (int i = 0; i < 9; ++i) sig[i] = 2.0 * mu * c.Es[i] + ((i % 4 == 0) ? lam * trE : T(0));
    mm3(R, sig, c.Rsig);
    // G = (tr S I - S) R^T; the reference's invertibility test uses |det G| > 1e-12 (Eigen default)
    T trS = c.S[0] + c.S[4] + c.S[8];
    T G0[9];
    for (int i = 0; i < 9; ++i) G0[i] = ((i % 4 == 0) ? trS : T(0)) - c.S[i];
    T detG = inv3(G0, c.G0i) * detR;
    c.correction = true;
    if (fabs(detR - 1.0) > 1e-8) {
        c.flags |= ELEM_FLAG_DET_R;
        c.correction = false;
    } else if (!(fabs(detG) > 1e-12)) {
        c.flags |= ELEM_FLAG_G_SINGULAR;
        c.correction = false;
    }
}

// f_a = -V R sigma g_a (doUpdateFMDK :101-113), ga = g_a
template <typename T>
__device__ __forceinline__ void coro_force(const CoroCommon<T>& c, const double* gav, double V, T* fa) {
    T t[3];
    const T ga[3] = {T(gav[0]), T(gav[1]), T(gav[2])};
    mv3(c.Rsig, ga, t);
    fa[0] = -V * t[0];
    fa[1] = -V * t[1];
    fa[2] = -V * t[2];
}

// Correction factor W_a of node a (zero when the correction is dropped).
// Quirk term: H' = R^T (S - I) = R^T E; p_a = V [lambda tr(H') g_a + mu (H' + H'^T) g_a]; q_a = R p_a;
// N_a = -[q]x + V mu ([S g]x - S [g]x); W_a = R N_a G^-1.
template <typename T>
__device__ __forceinline__ void coro_weight(const CoroCommon<T>& c, const double* ga_, double V, double lam,
                                            double mu, T* W) {
    if (!c.correction) {
        for (int i = 0; i < 9; ++i) W[i] = T(0);
        return;
    }
    const T* R = c.R;
    T Hp[9];
    mtm3(R, c.Es, Hp);
    T trHp = Hp[0] + Hp[4] + Hp[8];
    T p[3];
    for (int i = 0; i < 3; ++i) {
        T s = lam * trHp * ga_[i];
        for (int k = 0; k < 3; ++k) s += mu * (Hp[3 * i + k] + Hp[3 * k + i]) * ga_[k];
        p[i] = V * s;
    }
    T q[3];
    mv3(R, p, q);
    T Sg[3];
    const T ga[3] = {T(ga_[0]), T(ga_[1]), T(ga_[2])};
    mv3(c.S, ga, Sg);
    T N[9];
    // [w]x = [[0,-w2,w1],[w2,0,-w0],[-w1,w0,0]]
    double vm = V * mu;
    T sx[9] = {T(0), -Sg[2], Sg[1], Sg[2], T(0), -Sg[0], -Sg[1], Sg[0], T(0)};
    T qx[9] = {T(0), -q[2], q[1], q[2], T(0), -q[0], -q[1], q[0], T(0)};
    T gx[9] = {T(0), T(-ga_[2]), T(ga_[1]), T(ga_[2]), T(0), T(-ga_[0]), T(-ga_[1]), T(ga_[0]), T(0)};
    T Sgx[9];
    mm3(c.S, gx, Sgx);
    for (int i = 0; i < 9; ++i) N[i] = -qx[i] + vm * (sx[i] - Sgx[i]);
    T RN[9];
    mm3(R, N, RN);
    mm3(RN, c.G0i, W);
}

// Rotation, force and stiffness factors of one corotational tet at positions x (12) with rest x0 (12),
// rest gradients g and signed rest volume V. T is the arithmetic, TK the factor storage. Force f (12).
// Returns element flags.
template <typename TK, typename T = double>
__device__ inline int coro_factors(const double* x0, const T* x, const double g[4][3], double V, double lam,
                                   double mu, T* f, CoroFactors<TK>& cf) {
    CoroCommon<T> c;
    coro_common<T>(x0, x, g, lam, mu, c);
    for (int a = 0; a < 4; ++a) coro_force<T>(c, g[a], V, f + 3 * a);
    cf.correction = c.correction;
    for (int i = 0; i < 9; ++i) cf.R[i] = (TK)c.R[i];
    cf.V = (TK)V;
    cf.lam = (TK)lam;
    cf.mu = (TK)mu;
    for (int a = 0; a < 4; ++a) {
        T h[3];
        const T ga[3] = {T(g[a][0]), T(g[a][1]), T(g[a][2])};
        mv3(c.R, ga, h);
        for (int k = 0; k < 3; ++k) {
            cf.h[a][k] = (TK)h[k];
            cf.g[a][k] = (TK)g[a][k];
        }
    }
    for (int a = 0; a < 4; ++a) {
        T W[9];
        coro_weight<T>(c, g[a], V, lam, mu, W);
        for (int i = 0; i < 9; ++i) cf.W[a][i] = (TK)W[i];
    }
    return c.flags;
}

// K_ab from one row's factors: ha = R g_a, hb = R g_b, ga, gb, W = W_a, R; the same arithmetic as
// coro_block.
__device__ __forceinline__ void coro_row_block(const double* ha, const double* hb, const double* ga, const double* gb,
                                               const double* W, const double* R, bool correction, double V,
                                               double lam, double mu, double* K) {
    double gab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            K[3 * i + j] = V * (lam * ha[i] * hb[j] + mu * hb[i] * ha[j] + ((i == j) ? mu * gab : 0.0));
    if (correction) {
        double WG[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            WG[3 * i + 0] = W[3 * i + 1] * gb[2] - W[3 * i + 2] * gb[1];
            WG[3 * i + 1] = -W[3 * i + 0] * gb[2] + W[3 * i + 2] * gb[0];
            WG[3 * i + 2] = W[3 * i + 0] * gb[1] - W[3 * i + 1] * gb[0];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                K[3 * i + j] += WG[3 * i] * R[3 * j] + WG[3 * i + 1] * R[3 * j + 1] + WG[3 * i + 2] * R[3 * j + 2];
    }
}

// Linear tet stiffness block from the two gradients: K0_ab = V [lambda g_a g_b^T + mu g_b g_a^T +
// mu (g_a . g_b) I]; the same arithmetic as linear_block.
__device__ __forceinline__ void linear_row_block(const double* ga, const double* gb, double V, double lam, double mu,
                                                 double* K) {
    double gab = dot3(ga, gb);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K[3 * i + j] = V * (lam * ga[i] * gb[j] + mu * gb[i] * ga[j] + ((i == j) ? mu * gab : 0.0));
}

// Stiffness block K_ab (row-major 3x3) from the factors.
template <typename TK>
__device__ __forceinline__ void coro_block(const CoroFactors<TK>& cf, int a, int b, TK* K) {
    const TK* ha = cf.h[a];
    const TK* hb = cf.h[b];
    TK gab = cf.g[a][0] * cf.g[b][0] + cf.g[a][1] * cf.g[b][1] + cf.g[a][2] * cf.g[b][2];
    TK V = cf.V, lam = cf.lam, mu = cf.mu;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            K[3 * i + j] = V * (lam * ha[i] * hb[j] + mu * hb[i] * ha[j] + ((i == j) ? mu * gab : TK(0)));
    if (cf.correction) {
        // C_ab = W_a [g_b]x R^T
        const TK* W = cf.W[a];
        const TK* gb = cf.g[b];
        TK WG[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            // (W [g]x)(i, :) = W(i,:) [g]x ; [g]x columns: (0,g2,-g1), (-g2,0,g0), (g1,-g0,0)
            WG[3 * i + 0] = W[3 * i + 1] * gb[2] - W[3 * i + 2] * gb[1];
            WG[3 * i + 1] = -W[3 * i + 0] * gb[2] + W[3 * i + 2] * gb[0];
            WG[3 * i + 2] = W[3 * i + 0] * gb[1] - W[3 * i + 1] * gb[0];
        }
        const TK* R = cf.R;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                K[3 * i + j] += WG[3 * i] * R[3 * j] + WG[3 * i + 1] * R[3 * j + 1] + WG[3 * i + 2] * R[3 * j + 2];
    }
}

// Linear tet stiffness block K0_ab = V [lambda g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I]
// (equal to the block of V B^T E B, Physics/Fem3DElementTetrahedron.cpp:168-216).
template <typename TK>
__device__ __forceinline__ void linear_block(const double g[4][3], double V, double lam, double mu, int a, int b,
                                             TK* K) {
    double gab = dot3(g[a], g[b]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            K[3 * i + j] = (TK)(V * (lam * g[a][i] * g[b][j] + mu * g[b][i] * g[a][j] + ((i == j) ? mu * gab : 0.0)));
}

}  // namespace femgpu
