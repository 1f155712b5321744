// FP64 CPU oracle for the OpenSurgSim Fem3D implicit-step path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++17 without Eigen, the reference
// algorithm of the path (element kernels, sparsity pattern, assembly, boundary conditions,
// Eigen-ConjugateGradient semantics, implicit Euler). It is linked only by tests/, by the
// smoke check and by the CPU-baseline leg of bench.py, where it is the checker; it is never
// part of the shipped library or of its compute path.
//
// Every function cites the reference file:line it follows (paths relative to the reference
// root, SurgSim/...). Functions whose semantics come from Eigen (not vendored by the
// reference) restate Eigen 3.3's published algorithm and say so.
//
// Parity notes (what pins each function) are in the test files that exercise it.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// ------------------------------------------------------------------------------------------
// Small dense helpers (row-major fixed arrays).
// ------------------------------------------------------------------------------------------
struct M3 {
    double m[3][3];
};

static inline M3 m3_zero() { M3 r; std::memset(&r, 0, sizeof r); return r; }
static inline M3 m3_eye() { M3 r = m3_zero(); r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0; return r; }
static inline M3 m3_mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
static inline M3 m3_t(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
static inline double m3_det(const M3& a) {
    return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
           a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
           a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}

// Eigen's computeInverseAndDetWithCheck for 3x3 (cofactor inverse; invertible iff
// |det| > threshold, threshold defaulting to NumTraits<double>::dummy_precision() = 1e-12).
// Used at Physics/Fem3DElementCorotationalTetrahedron.cpp:210 and Physics/Fem3DElementCube.cpp:202.
static bool m3_inverse_with_check(const M3& a, M3* inv, double* det, double threshold = 1e-12) {
    double c00 = a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1];
    double c01 = a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2];
    double c02 = a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0];
    double d = a.m[0][0] * c00 + a.m[0][1] * c01 + a.m[0][2] * c02;
    *det = d;
    if (!(std::fabs(d) > threshold)) return false;
    double id = 1.0 / d;
    inv->m[0][0] = c00 * id;
    inv->m[1][0] = c01 * id;
    inv->m[2][0] = c02 * id;
    inv->m[0][1] = (a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2]) * id;
    inv->m[1][1] = (a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0]) * id;
    inv->m[2][1] = (a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1]) * id;
    inv->m[0][2] = (a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]) * id;
    inv->m[1][2] = (a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2]) * id;
    inv->m[2][2] = (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]) * id;
    return true;
}

// 4x4 inverse with determinant check (Gauss-Jordan with partial pivoting; invertible iff
// |det| > 1e-12, the Eigen default). Restates V.computeInverseAndDetWithCheck at
// Physics/Fem3DElementCorotationalTetrahedron.cpp:57-72.
static bool m4_inverse_with_check(const double a_in[4][4], double inv[4][4], double* det_out) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            a[i][j] = a_in[i][j];
            a[i][j + 4] = (i == j) ? 1.0 : 0.0;
        }
    double det = 1.0;
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) {
            *det_out = 0.0;
            return false;
        }
        if (p != c) {
            for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
            det = -det;
        }
        det *= a[c][c];
        double ip = 1.0 / a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] *= ip;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
        }
    }
    *det_out = det;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) inv[i][j] = a[i][j + 4];
    return std::fabs(det) > 1e-12;
}

// SurgSim::Math::skew (Math/Matrix.h:112-117): vector of the skew-symmetric part of a matrix.
static inline void skew_vec(const M3& a, double v[3]) {
    double s21 = (a.m[2][1] - a.m[1][2]) / 2.0;
    double s02 = (a.m[0][2] - a.m[2][0]) / 2.0;
    double s10 = (a.m[1][0] - a.m[0][1]) / 2.0;
    v[0] = s21;
    v[1] = s02;
    v[2] = s10;
}
// SurgSim::Math::makeSkewSymmetricMatrix (Math/Matrix.h:85-104).
static inline M3 skew_mat(const double v[3]) {
    M3 r;
    r.m[0][0] = 0.0;   r.m[0][1] = -v[2]; r.m[0][2] = v[1];
    r.m[1][0] = v[2];  r.m[1][1] = 0.0;   r.m[1][2] = -v[0];
    r.m[2][0] = -v[1]; r.m[2][1] = v[0];  r.m[2][2] = 0.0;
    return r;
}

// ------------------------------------------------------------------------------------------
// 3x3 SVD and Eigen's Transform::computeRotationScaling.
//
// Eigen (not vendored; >= 3.3) defines computeRotationScaling on the JacobiSVD of the linear
// part: x = det(U V^T); sv(0) *= x; S = V diag(sv) V^T; R = U diag(1/x, 1, 1) V^T, with
// singular values sorted in decreasing order and non-negative. Called at
// Physics/Fem3DElementCorotationalTetrahedron.cpp:146-154. The SVD here is a one-sided
// (Hestenes) Jacobi iteration; U, V are unique up to simultaneous column signs for distinct
// singular values, which leaves R and S unchanged, so any accurate SVD gives the same R, S.
// ------------------------------------------------------------------------------------------
static void svd3(const M3& F, M3* U, double s[3], M3* V) {
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = F.m[i][j];
    M3 W = m3_eye();
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < 3; ++i) {
                    alpha += A[i][p] * A[i][p];
                    beta += A[i][q] * A[i][q];
                    gamma += A[i][p] * A[i][q];
                }
                if (gamma == 0.0 || std::fabs(gamma) <= 1e-300 ||
                    std::fabs(gamma) <= 1e-17 * std::sqrt(alpha * beta))
                    continue;
                rotated = true;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t);
                double sn = c * t;
                for (int i = 0; i < 3; ++i) {
                    double ap = A[i][p], aq = A[i][q];
                    A[i][p] = c * ap - sn * aq;
                    A[i][q] = sn * ap + c * aq;
                    double wp = W.m[i][p], wq = W.m[i][q];
                    W.m[i][p] = c * wp - sn * wq;
                    W.m[i][q] = sn * wp + c * wq;
                }
            }
        if (!rotated) break;
    }
    double sig[3];
    for (int j = 0; j < 3; ++j) sig[j] = std::sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int a, int b) { return sig[a] > sig[b]; });
    M3 Uo = m3_zero(), Vo = m3_zero();
    for (int k = 0; k < 3; ++k) {
        int j = order[k];
        s[k] = sig[j];
        for (int i = 0; i < 3; ++i) {
            Vo.m[i][k] = W.m[i][j];
            Uo.m[i][k] = (sig[j] > 0) ? A[i][j] / sig[j] : 0.0;
        }
    }
    // Complete U where singular values vanish (rank-deficient F).
    auto col = [&](M3& M, int k, double v[3]) { for (int i = 0; i < 3; ++i) v[i] = M.m[i][k]; };
    auto setcol = [&](M3& M, int k, const double v[3]) { for (int i = 0; i < 3; ++i) M.m[i][k] = v[i]; };
    if (!(s[2] > 0)) {
        if (!(s[1] > 0)) {
            double u0[3];
            col(Uo, 0, u0);
            if (!(s[0] > 0)) { u0[0] = 1; u0[1] = 0; u0[2] = 0; setcol(Uo, 0, u0); }
            double e[3] = {std::fabs(u0[0]) < 0.9 ? 1.0 : 0.0, std::fabs(u0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
            double d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
            double u1[3] = {e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]};
            double n = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
            for (int i = 0; i < 3; ++i) u1[i] /= n;
            setcol(Uo, 1, u1);
        }
        double u0[3], u1[3];
        col(Uo, 0, u0);
        col(Uo, 1, u1);
        double u2[3] = {u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2], u0[0] * u1[1] - u0[1] * u1[0]};
        setcol(Uo, 2, u2);
    }
    *U = Uo;
    *V = Vo;
}

static void rotation_scaling(const M3& F, M3* R, M3* S) {
    M3 U, V;
    double sv[3];
    svd3(F, &U, sv, &V);
    double x = m3_det(m3_mul(U, m3_t(V)));
    sv[0] *= x;
    M3 Sm = m3_zero(), Rm = m3_zero();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += V.m[i][k] * sv[k] * V.m[j][k];
            Sm.m[i][j] = s;
        }
    M3 Ux = U;
    for (int i = 0; i < 3; ++i) Ux.m[i][0] /= x;
    Rm = m3_mul(Ux, m3_t(V));
    *R = Rm;
    *S = Sm;
}

// ------------------------------------------------------------------------------------------
// Linear tetrahedron (Physics/Fem3DElementTetrahedron.cpp).
// ------------------------------------------------------------------------------------------
// det() helper, Physics/Fem3DElementTetrahedron.cpp:34-38.
static inline double det3v(const double a[3], const double b[3], const double c[3]) {
    return a[0] * b[1] * c[2] + a[2] * b[0] * c[1] + a[1] * b[2] * c[0] - a[2] * b[1] * c[0] -
           a[1] * b[0] * c[2] - a[0] * b[2] * c[1];
}

// getVolume, Physics/Fem3DElementTetrahedron.cpp:218-235 (signed).
static double tet_volume(const double* p0, const double* p1, const double* p2, const double* p3) {
    return (det3v(p1, p2, p3) - det3v(p0, p2, p3) + det3v(p0, p1, p3) - det3v(p0, p1, p2)) / 6.0;
}

// computeShapeFunctions, Physics/Fem3DElementTetrahedron.cpp:237-359.
static void tet_shape_functions(const double x0[12], double* vol, double ai[4], double bi[4], double ci[4],
                                double di[4]) {
    const double* a = x0;
    const double* b = x0 + 3;
    const double* c = x0 + 6;
    const double* d = x0 + 9;
    *vol = tet_volume(a, b, c, d);
    ai[0] = det3v(b, c, d);
    ai[1] = -det3v(a, c, d);
    ai[2] = det3v(a, b, d);
    ai[3] = -det3v(a, b, c);
    {
        double at[3] = {1, a[1], a[2]}, bt[3] = {1, b[1], b[2]}, ct[3] = {1, c[1], c[2]}, dt[3] = {1, d[1], d[2]};
        bi[0] = -det3v(bt, ct, dt);
        bi[1] = det3v(at, ct, dt);
        bi[2] = -det3v(at, bt, dt);
        bi[3] = det3v(at, bt, ct);
    }
    {
        double at[3] = {1, a[0], a[2]}, bt[3] = {1, b[0], b[2]}, ct[3] = {1, c[0], c[2]}, dt[3] = {1, d[0], d[2]};
        ci[0] = det3v(bt, ct, dt);
        ci[1] = -det3v(at, ct, dt);
        ci[2] = det3v(at, bt, dt);
        ci[3] = -det3v(at, bt, ct);
    }
    {
        double at[3] = {1, a[0], a[1]}, bt[3] = {1, b[0], b[1]}, ct[3] = {1, c[0], c[1]}, dt[3] = {1, d[0], d[1]};
        di[0] = -det3v(bt, ct, dt);
        di[1] = det3v(at, ct, dt);
        di[2] = -det3v(at, bt, dt);
        di[3] = det3v(at, bt, ct);
    }
}

// Material asserts of FemElement::initialize, Physics/FemElement.cpp:33-43.
static void check_material(double E, double nu, double rho) {
    if (rho == 0.0) throw std::runtime_error("Mass density is not set");
    if (nu == 0.0) throw std::runtime_error("Poisson ratio is not set");
    if (E == 0.0) throw std::runtime_error("Young modulus is not set");
    if (!(rho > 0.0)) throw std::runtime_error("Mass density is invalid, it should be positive");
    if (!(nu > 0.0 && nu < 0.5)) throw std::runtime_error("Poisson ratio is invalid, it should be within [0 0.5)");
    if (!(E > 0.0)) throw std::runtime_error("Young modulus is invalid, it should be positive");
}

// computeStiffness, Physics/Fem3DElementTetrahedron.cpp:168-216 (K = V B^T E B, symmetrized).
static void tet_linear_stiffness(const double x0[12], double E, double nu, double K[144], double* vol_out) {
    double vol, ai[4], bi[4], ci[4], di[4];
    tet_shape_functions(x0, &vol, ai, bi, ci, di);
    double B[6][12];
    std::memset(B, 0, sizeof B);
    double coef = 1.0 / (6.0 * vol);
    for (int i = 0; i < 4; ++i) {
        B[0][3 * i] = coef * bi[i];
        B[1][3 * i + 1] = coef * ci[i];
        B[2][3 * i + 2] = coef * di[i];
        B[3][3 * i] = coef * ci[i];
        B[3][3 * i + 1] = coef * bi[i];
        B[4][3 * i + 1] = coef * di[i];
        B[4][3 * i + 2] = coef * ci[i];
        B[5][3 * i] = coef * di[i];
        B[5][3 * i + 2] = coef * bi[i];
    }
    double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1 + nu));
    double Em[6][6];
    std::memset(Em, 0, sizeof Em);
    Em[0][0] = Em[1][1] = Em[2][2] = 2.0 * mu + lambda;
    Em[0][1] = Em[0][2] = Em[1][0] = Em[1][2] = Em[2][0] = Em[2][1] = lambda;
    Em[3][3] = Em[4][4] = Em[5][5] = mu;
    double EB[6][12];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += Em[i][k] * B[k][j];
            EB[i][j] = s;
        }
    double Kt[12][12];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += B[k][i] * EB[k][j];
            Kt[i][j] = vol * s;
        }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) K[i * 12 + j] = (Kt[i][j] + Kt[j][i]) * 0.5;
    if (vol_out) *vol_out = vol;
}

// computeMass, Physics/Fem3DElementTetrahedron.cpp:123-166 (rho V / 20 (2 diag, 1 off) x I3).
static void tet_linear_mass(double vol, double rho, double M[144]) {
    double coef = vol * rho / 20.0;
    std::memset(M, 0, 144 * sizeof(double));
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 3; ++c) M[(3 * a + c) * 12 + 3 * b + c] = (a == b) ? 2.0 * coef : coef;
}

// ------------------------------------------------------------------------------------------
// Element types.
// ------------------------------------------------------------------------------------------
enum ElementKind { TET_LINEAR = 0, TET_COROTATIONAL = 1, HEX = 2 };

// Counters for the reference's degenerate branches.
struct Counters {
    long det_r_invalid = 0;   // Physics/Fem3DElementCorotationalTetrahedron.cpp:161-166
    long g_singular = 0;      // Physics/Fem3DElementCorotationalTetrahedron.cpp:211-216
};

struct Element {
    int kind = TET_LINEAR;
    int nn = 4;
    int ids[8];
    double E = 0, nu = 0, rho = 0;
    std::vector<double> x0;        // rest positions (3 nn)
    std::vector<double> Klin, Mlin; // linear K, M (corotational: m_KLinear / m_MLinear)
    std::vector<double> K, M, f;    // current m_K, m_M, m_f
    double Vinv[4][4];
    M3 R = m3_eye();
    double volume = 0;
};

// Fem3DElementCorotationalTetrahedron::computeRotationMassAndStiffness, literal restatement of
// Physics/Fem3DElementCorotationalTetrahedron.cpp:116-288 including the in-place use of m_K at
// :276 (Ke == &m_K on the FMDK path, :83-97). The dense 12x12 products are kept as written
// (the reference's own cost model).
static void coro_rotation_mass_stiffness(Element& e, const double x[12], bool updR, bool updM, bool updK,
                                         Counters* cnt, const M3* R_given = nullptr) {
    double P[4][4];
    for (int a = 0; a < 4; ++a) {
        for (int c = 0; c < 3; ++c) P[c][a] = x[3 * a + c];
        P[3][a] = 1.0;
    }
    double PV[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += P[i][k] * e.Vinv[k][j];
            PV[i][j] = s;
        }
    M3 F;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F.m[i][j] = PV[i][j];
    M3 R, S;
    if (R_given) {
        // A caller-supplied polar factor (F = R S): S = sym(R^T F). Used where computeRotationScaling's R is
        // not unique (tied singular values), to check everything downstream of R against another R.
        R = *R_given;
        const M3 RtF = m3_mul(m3_t(R), F);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S.m[i][j] = 0.5 * (RtF.m[i][j] + RtF.m[j][i]);
    } else {
        rotation_scaling(F, &R, &S);
    }
    if (updR) e.R = R;
    if (std::fabs(m3_det(R) - 1.0) > 1e-8) {
        if (cnt) cnt->det_r_invalid++;
        return;  // M and K left stale (:161-166)
    }
    double R12[12][12];
    std::memset(R12, 0, sizeof R12);
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R12[3 * a + i][3 * a + j] = R.m[i][j];
    if (updM) {
        double RM[12][12];
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += R12[i][k] * e.Mlin[k * 12 + j];
                RM[i][j] = s;
            }
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += RM[i][k] * R12[j][k];
                e.M[i * 12 + j] = s;
            }
    }
    if (!updK) return;
    double RK[12][12];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0.0;
            for (int k = 0; k < 12; ++k) s += R12[i][k] * e.Klin[k * 12 + j];
            RK[i][j] = s;
        }
    // *Ke = RK * R12^T  (:203) -- overwrites m_K.
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0.0;
            for (int k = 0; k < 12; ++k) s += RK[i][k] * R12[j][k];
            e.K[i * 12 + j] = s;
        }
    // G = (tr(S) I - S) R^T (:208)
    double trS = S.m[0][0] + S.m[1][1] + S.m[2][2];
    M3 TS;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) TS.m[i][j] = (i == j ? trS : 0.0) - S.m[i][j];
    M3 G = m3_mul(TS, m3_t(R));
    M3 Ginv = m3_zero();
    double detG;
    if (!m3_inverse_with_check(G, &Ginv, &detG)) {
        if (cnt) cnt->g_singular++;
        return;  // K = R K0 R^T only (:211-216)
    }
    // dR/dF_ij = [w_ij] R with G w_ij = 2 skew(R^T e_i e_j^T)   (:220-233)
    M3 dRdF[3][3];
    M3 Rt = m3_t(R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            M3 Eij = m3_zero();
            Eij.m[i][j] = 1.0;
            M3 RtE = m3_mul(Rt, Eij);
            double sv[3];
            skew_vec(RtE, sv);
            double w[3];
            for (int r = 0; r < 3; ++r) {
                w[r] = 0.0;
                for (int k = 0; k < 3; ++k) w[r] += Ginv.m[r][k] * 2.0 * sv[k];
            }
            dRdF[i][j] = m3_mul(skew_mat(w), R);
        }
    // dR/dx (:251-269)
    M3 dRdX[12];
    for (int node = 0; node < 4; ++node) {
        double ni[3] = {e.Vinv[node][0], e.Vinv[node][1], e.Vinv[node][2]};
        for (int axis = 0; axis < 3; ++axis) {
            int dof = 3 * node + axis;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    dRdX[dof].m[i][j] = dRdF[i][j].m[axis][0] * ni[0] + dRdF[i][j].m[axis][1] * ni[1] +
                                        dRdF[i][j].m[axis][2] * ni[2];
        }
    }
    // KTimesRx_x0 = m_K * (R12^T x - x0)   (:276)   m_K == R K0 R^T here (quirk)
    double Rtx_x0[12];
    for (int i = 0; i < 12; ++i) {
        double s = 0.0;
        for (int k = 0; k < 12; ++k) s += R12[k][i] * x[k];
        Rtx_x0[i] = s - e.x0[i];
    }
    double KT[12];
    for (int i = 0; i < 12; ++i) {
        double s = 0.0;
        for (int k = 0; k < 12; ++k) s += e.K[i * 12 + k] * Rtx_x0[k];
        KT[i] = s;
    }
    // (:277-286)
    for (int dof = 0; dof < 12; ++dof) {
        double D12[12][12];
        std::memset(D12, 0, sizeof D12);
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) D12[3 * a + i][3 * a + j] = dRdX[dof].m[i][j];
        // RK * D12^T
        double RKDt[12][12];
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += RK[i][k] * D12[j][k];
                RKDt[i][j] = s;
            }
        for (int i = 0; i < 12; ++i) {
            double s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < 12; ++k) {
                s1 += D12[i][k] * KT[k];
                s2 += RKDt[i][k] * x[k];
            }
            e.K[i * 12 + dof] += s1 + s2;
        }
    }
}

// Cube element helpers, Physics/Fem3DElementCube.cpp.
static const double kEpsSign[8] = {-1.0, +1.0, +1.0, -1.0, -1.0, +1.0, +1.0, -1.0};  // :72
static const double kEtaSign[8] = {-1.0, -1.0, +1.0, +1.0, -1.0, -1.0, +1.0, +1.0};  // :73
static const double kMuSign[8] = {-1.0, -1.0, -1.0, -1.0, +1.0, +1.0, +1.0, +1.0};   // :74
// Math/GaussLegendreQuadrature.cpp:30-34 (order +1/sqrt3, -1/sqrt3; weights 1).
static const double kGauss2[2] = {1.0 / 1.7320508075688772935274463415059, -1.0 / 1.7320508075688772935274463415059};

static inline double cube_N(int i, double e, double n, double m) {
    return 1.0 / 8.0 * (1 + e * kEpsSign[i]) * (1 + n * kEtaSign[i]) * (1 + m * kMuSign[i]);
}
static inline double cube_dNde(int i, double e, double n, double m) {
    (void)e;
    return 1.0 / 8.0 * kEpsSign[i] * (1 + n * kEtaSign[i]) * (1 + m * kMuSign[i]);
}
static inline double cube_dNdn(int i, double e, double n, double m) {
    (void)n;
    return 1.0 / 8.0 * (1 + e * kEpsSign[i]) * kEtaSign[i] * (1 + m * kMuSign[i]);
}
static inline double cube_dNdm(int i, double e, double n, double m) {
    (void)m;
    return 1.0 / 8.0 * (1 + e * kEpsSign[i]) * (1 + n * kEtaSign[i]) * kMuSign[i];
}

// evaluateJ, Physics/Fem3DElementCube.cpp:169-211.
static void cube_J(const double* p, double e, double n, double m, M3* J, M3* Jinv, double* detJ) {
    *J = m3_zero();
    for (int idx = 0; idx < 8; ++idx)
        for (int axis = 0; axis < 3; ++axis) {
            J->m[0][axis] += p[3 * idx + axis] * cube_dNde(idx, e, n, m);
            J->m[1][axis] += p[3 * idx + axis] * cube_dNdn(idx, e, n, m);
            J->m[2][axis] += p[3 * idx + axis] * cube_dNdm(idx, e, n, m);
        }
    if (Jinv && detJ) {
        if (!m3_inverse_with_check(*J, Jinv, detJ))
            throw std::runtime_error("Found a non invertible matrix J while computing Fem3DElementCube stiffness matrix");
    }
}

// getVolume, Physics/Fem3DElementCube.cpp:321-366.
static double cube_volume(const double* p) {
    double v = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                M3 J, Jinv;
                double detJ;
                cube_J(p, kGauss2[i], kGauss2[j], kGauss2[k], &J, &Jinv, &detJ);
                v += 1.0 * 1.0 * 1.0 * detJ;
            }
    if (!(v > 1e-12)) throw std::runtime_error("Fem3DElementCube ill-defined, its volume is not positive");
    return v;
}

// computeMass / addMassMatrixAtPoint, Physics/Fem3DElementCube.cpp:97-127, 290-319.
static void cube_mass(const double* p, double rho, double* M) {
    std::memset(M, 0, 576 * sizeof(double));
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double e = kGauss2[i], n = kGauss2[j], m = kGauss2[k];
                M3 J, Jinv;
                double detJ;
                cube_J(p, e, n, m, &J, &Jinv, &detJ);
                double phi[8];
                for (int idx = 0; idx < 8; ++idx) phi[idx] = cube_N(idx, e, n, m);
                double w = 1.0 * 1.0 * 1.0 * detJ * rho;
                for (int a = 0; a < 8; ++a)
                    for (int b = 0; b < 8; ++b)
                        for (int c = 0; c < 3; ++c) M[(3 * a + c) * 24 + 3 * b + c] += w * (phi[a] * phi[b]);
            }
}

// computeStiffness / evaluateStrainDisplacement / addStrainStressStiffnessAtPoint,
// Physics/Fem3DElementCube.cpp:143-167, 213-288.
static void cube_stiffness(const double* p, double E, double nu, double* K) {
    std::memset(K, 0, 576 * sizeof(double));
    double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1 + nu));
    double C[6][6];
    std::memset(C, 0, sizeof C);
    C[0][0] = C[1][1] = C[2][2] = 2.0 * mu + lambda;
    C[0][1] = C[0][2] = C[1][0] = lambda;
    C[1][2] = C[2][0] = C[2][1] = lambda;
    C[3][3] = C[4][4] = C[5][5] = mu;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double e = kGauss2[i], n = kGauss2[j], m = kGauss2[k];
                M3 J, Jinv;
                double detJ;
                cube_J(p, e, n, m, &J, &Jinv, &detJ);
                double B[6][24];
                std::memset(B, 0, sizeof B);
                for (int idx = 0; idx < 8; ++idx) {
                    double dN[3] = {cube_dNde(idx, e, n, m), cube_dNdn(idx, e, n, m), cube_dNdm(idx, e, n, m)};
                    double g[3];
                    for (int r = 0; r < 3; ++r) g[r] = Jinv.m[r][0] * dN[0] + Jinv.m[r][1] * dN[1] + Jinv.m[r][2] * dN[2];
                    B[0][3 * idx] = g[0];
                    B[1][3 * idx + 1] = g[1];
                    B[2][3 * idx + 2] = g[2];
                    B[3][3 * idx] = g[1];
                    B[3][3 * idx + 1] = g[0];
                    B[4][3 * idx + 1] = g[2];
                    B[4][3 * idx + 2] = g[1];
                    B[5][3 * idx] = g[2];
                    B[5][3 * idx + 2] = g[0];
                }
                double w = 1.0 * 1.0 * 1.0 * detJ;
                double CB[6][24];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 24; ++c) {
                        double s = 0.0;
                        for (int q = 0; q < 6; ++q) s += C[r][q] * B[q][c];
                        CB[r][c] = s;
                    }
                for (int r = 0; r < 24; ++r)
                    for (int c = 0; c < 24; ++c) {
                        double s = 0.0;
                        for (int q = 0; q < 6; ++q) s += B[q][r] * CB[q][c];
                        K[r * 24 + c] += w * s;
                    }
            }
}

static void element_init(Element& e, const double* xall, Counters* cnt) {
    check_material(e.E, e.nu, e.rho);
    e.x0.assign(3 * e.nn, 0.0);
    for (int a = 0; a < e.nn; ++a)
        for (int c = 0; c < 3; ++c) e.x0[3 * a + c] = xall[3 * (size_t)e.ids[a] + c];
    int n = 3 * e.nn;
    e.f.assign(n, 0.0);
    e.K.assign(n * n, 0.0);
    e.M.assign(n * n, 0.0);
    if (e.kind == HEX) {
        e.volume = cube_volume(e.x0.data());
        cube_mass(e.x0.data(), e.rho, e.M.data());
        cube_stiffness(e.x0.data(), e.E, e.nu, e.K.data());
        return;
    }
    // Physics/Fem3DElementTetrahedron.cpp:92-121
    double vol;
    tet_linear_stiffness(e.x0.data(), e.E, e.nu, e.K.data(), &vol);
    e.volume = vol;
    tet_linear_mass(vol, e.rho, e.M.data());
    if (e.kind == TET_COROTATIONAL) {
        // Physics/Fem3DElementCorotationalTetrahedron.cpp:48-75
        e.Mlin = e.M;
        e.Klin = e.K;
        double V[4][4];
        for (int a = 0; a < 4; ++a) {
            for (int c = 0; c < 3; ++c) V[c][a] = e.x0[3 * a + c];
            V[3][a] = 1.0;
        }
        double det;
        if (!m4_inverse_with_check(V, e.Vinv, &det))
            throw std::runtime_error("Trying to initialize an invalid co-rotational tetrahedron. Matrix V not invertible.");
        coro_rotation_mass_stiffness(e, e.x0.data(), true, true, true, cnt);
        // f at rest (doUpdateFMDK :101-113)
        double Rx0[12];
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i)
                Rx0[3 * a + i] = e.R.m[i][0] * e.x0[3 * a] + e.R.m[i][1] * e.x0[3 * a + 1] + e.R.m[i][2] * e.x0[3 * a + 2];
        for (int i = 0; i < 12; ++i) {
            double s = 0.0;
            for (int k = 0; k < 12; ++k) s += e.K[i * 12 + k] * (e.x0[k] - Rx0[k]);
            e.f[i] = -s;
        }
    }
}

// updateFMDK(state, FMDK) for one element (doUpdateFMDK of each type).
static void element_update(Element& e, const double* xall, int options, Counters* cnt,
                           const M3* R_given = nullptr) {
    int n = 3 * e.nn;
    double x[24];
    for (int a = 0; a < e.nn; ++a)
        for (int c = 0; c < 3; ++c) x[3 * a + c] = xall[3 * (size_t)e.ids[a] + c];
    const int F = 1, M = 2, K = 8;
    if (e.kind == TET_COROTATIONAL) {
        // Physics/Fem3DElementCorotationalTetrahedron.cpp:77-114
        bool updR = options & F;
        bool updK = (options & F) || (options & K);
        bool updM = options & M;
        coro_rotation_mass_stiffness(e, x, updR, updM, updK, cnt, R_given);
        if (options & F) {
            double Rx0[12];
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < 3; ++i)
                    Rx0[3 * a + i] = e.R.m[i][0] * e.x0[3 * a] + e.R.m[i][1] * e.x0[3 * a + 1] + e.R.m[i][2] * e.x0[3 * a + 2];
            for (int i = 0; i < 12; ++i) {
                double s = 0.0;
                for (int k = 0; k < 12; ++k) s += e.K[i * 12 + k] * (x[k] - Rx0[k]);
                e.f[i] = -s;
            }
        }
        return;
    }
    // Linear tet (Physics/Fem3DElementTetrahedron.cpp:77-90) and cube (Physics/Fem3DElementCube.cpp:129-141)
    if (options & F) {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += e.K[i * n + k] * (x[k] - e.x0[k]);
            e.f[i] = -s;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Sparse matrix (CSC, int32 indices; Eigen::SparseMatrix<double> ColMajor, Math/SparseMatrix.h:32).
// ------------------------------------------------------------------------------------------
struct Csc {
    int n = 0;
    std::vector<int> outer;  // n+1
    std::vector<int> inner;  // nnz
    std::vector<double> val; // nnz
    int find(int row, int col) const {
        auto b = inner.begin() + outer[col], e = inner.begin() + outer[col + 1];
        auto it = std::lower_bound(b, e, row);
        if (it == e || *it != row) return -1;
        return (int)(it - inner.begin());
    }
};

// Sparsity pattern: FemRepresentation::doInitialize, Physics/FemRepresentation.cpp:121-132.
// Every element inserts a full (3 nn)x(3 nn) zero block through assembleMatrixBlocks ->
// addSubMatrix -> blockOperation -> coeffRef (Physics/FemElement-inl.h:31-48,
// Math/SparseMatrix.h:266-302); makeCompressed() then yields, per column, the sorted row
// indices of all inserted coefficients.
static void build_pattern(int num_nodes, const std::vector<Element>& elems, Csc* A) {
    std::vector<std::vector<int>> nbr(num_nodes);
    for (const auto& e : elems)
        for (int a = 0; a < e.nn; ++a)
            for (int b = 0; b < e.nn; ++b) nbr[e.ids[b]].push_back(e.ids[a]);
    int n = 3 * num_nodes;
    A->n = n;
    A->outer.assign(n + 1, 0);
    A->inner.clear();
    for (int J = 0; J < num_nodes; ++J) {
        auto& v = nbr[J];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (int c = 0; c < 3; ++c) {
            int col = 3 * J + c;
            for (int I : v)
                for (int r = 0; r < 3; ++r) A->inner.push_back(3 * I + r);
            A->outer[col + 1] = (int)A->inner.size();
        }
    }
    A->val.assign(A->inner.size(), 0.0);
}

// assembleMatrixBlocksNoInitialize / blockWithSearch add: Physics/FemElement-inl.h:51-68,
// Math/SparseMatrix.h:178-242. Adds scale*Ke into the existing pattern (asserting presence).
static void add_element_matrix(const Element& e, const std::vector<double>& Ke, double scale, Csc* A) {
    int n = 3 * e.nn;
    for (int b = 0; b < e.nn; ++b)
        for (int cj = 0; cj < 3; ++cj) {
            int col = 3 * e.ids[b] + cj;
            for (int a = 0; a < e.nn; ++a) {
                int row0 = 3 * e.ids[a];
                int idx = A->find(row0, col);
                if (idx < 0 || idx + 2 >= A->outer[col + 1] || A->inner[idx + 2] != row0 + 2)
                    throw std::runtime_error("matrix is missing an element of the block");
                for (int ri = 0; ri < 3; ++ri) A->val[idx + ri] += scale * Ke[(3 * a + ri) * n + 3 * b + cj];
            }
        }
}

// y = A x (full matrix)
static void csc_matvec(const Csc& A, const double* x, double* y) {
    std::fill(y, y + A.n, 0.0);
    for (int j = 0; j < A.n; ++j) {
        double xj = x[j];
        for (int p = A.outer[j]; p < A.outer[j + 1]; ++p) y[A.inner[p]] += A.val[p] * xj;
    }
}

// y = selfadjointView<Lower>(A) x  (Eigen's sparse_selfadjoint_time_dense_product for a
// column-major matrix and Lower: diagonal once, strictly-lower entries used twice).
static void csc_symv_lower(const Csc& A, const double* x, double* y) {
    std::fill(y, y + A.n, 0.0);
    for (int j = 0; j < A.n; ++j) {
        int p = A.outer[j], pe = A.outer[j + 1];
        while (p < pe && A.inner[p] < j) ++p;
        double res_j = 0.0;
        double xj = x[j];
        if (p < pe && A.inner[p] == j) {
            y[j] += A.val[p] * xj;
            ++p;
        }
        for (; p < pe; ++p) {
            int i = A.inner[p];
            double v = A.val[p];
            res_j += v * x[i];
            y[i] += v * xj;
        }
        y[j] += res_j;
    }
}

struct CgResult {
    long iters = 0;
    double error = 0.0;
};

// Eigen::ConjugateGradient<SparseMatrix<double>> (defaults: UpLo=Lower,
// DiagonalPreconditioner) as used by LinearSparseSolveAndInverseCG
// (Math/LinearSparseSolveAndInverse.h:122, .cpp:119-129). Restates Eigen 3.3's
// internal::conjugate_gradient: x0 = 0 (solve() without guess), residual r = b - A x0,
// threshold = max(tol^2 |b|^2, realmin), recursive residual, maxIter default 2n,
// DiagonalPreconditioner::factorize (1/diag, 1 where the diagonal is absent or zero).
static CgResult cg_solve_lower(const Csc& A, const double* b, double* x, double tol, long max_iter) {
    int n = A.n;
    CgResult res;
    std::vector<double> invdiag(n, 1.0);
    for (int j = 0; j < n; ++j) {
        int idx = A.find(j, j);
        if (idx >= 0 && A.val[idx] != 0.0) invdiag[j] = 1.0 / A.val[idx];
    }
    if (max_iter < 0) max_iter = 2L * n;
    std::fill(x, x + n, 0.0);
    std::vector<double> r(b, b + n), p(n), z(n), tmp(n);
    // residual = rhs - mat * x  with x = 0
    double rhsNorm2 = 0.0;
    for (int i = 0; i < n; ++i) rhsNorm2 += b[i] * b[i];
    if (rhsNorm2 == 0.0) {
        res.iters = 0;
        res.error = 0.0;
        return res;
    }
    const double considerAsZero = std::numeric_limits<double>::min();
    double threshold = std::max(tol * tol * rhsNorm2, considerAsZero);
    double residualNorm2 = 0.0;
    for (int i = 0; i < n; ++i) residualNorm2 += r[i] * r[i];
    if (residualNorm2 < threshold) {
        res.iters = 0;
        res.error = std::sqrt(residualNorm2 / rhsNorm2);
        return res;
    }
    for (int i = 0; i < n; ++i) p[i] = invdiag[i] * r[i];
    double absNew = 0.0;
    for (int i = 0; i < n; ++i) absNew += r[i] * p[i];
    long i = 0;
    while (i < max_iter) {
        csc_symv_lower(A, p.data(), tmp.data());
        double pAp = 0.0;
        for (int k = 0; k < n; ++k) pAp += p[k] * tmp[k];
        double alpha = absNew / pAp;
        for (int k = 0; k < n; ++k) x[k] += alpha * p[k];
        for (int k = 0; k < n; ++k) r[k] -= alpha * tmp[k];
        residualNorm2 = 0.0;
        for (int k = 0; k < n; ++k) residualNorm2 += r[k] * r[k];
        if (residualNorm2 < threshold) break;
        for (int k = 0; k < n; ++k) z[k] = invdiag[k] * r[k];
        double absOld = absNew;
        absNew = 0.0;
        for (int k = 0; k < n; ++k) absNew += r[k] * z[k];
        double beta = absNew / absOld;
        for (int k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
        i++;
    }
    res.error = std::sqrt(residualNorm2 / rhsNorm2);
    res.iters = i;
    return res;
}

// Direct solve for LinearSparseSolveAndInverseLU (Eigen SparseLU, .cpp:31-97): banded LU
// without pivoting over the full matrix. SparseLU's ordering and pivoting change only rounding
// for the diagonally dominant systems of this path.
struct BandLU {
    int n = 0, bw = 0;
    std::vector<double> a;  // (2 bw + 1) x n band storage: a[(i - j + bw) * n + j]
    bool ok = false;
    double& at(int i, int j) { return a[(size_t)(i - j + bw) * n + j]; }
    void factor(const Csc& A) {
        n = A.n;
        bw = 0;
        for (int j = 0; j < n; ++j)
            for (int p = A.outer[j]; p < A.outer[j + 1]; ++p) bw = std::max(bw, std::abs(A.inner[p] - j));
        a.assign((size_t)(2 * bw + 1) * n, 0.0);
        for (int j = 0; j < n; ++j)
            for (int p = A.outer[j]; p < A.outer[j + 1]; ++p) at(A.inner[p], j) = A.val[p];
        for (int k = 0; k < n; ++k) {
            double piv = at(k, k);
            if (piv == 0.0) throw std::runtime_error("LU factorization failed: zero pivot");
            int iend = std::min(n - 1, k + bw);
            for (int i = k + 1; i <= iend; ++i) {
                double l = at(i, k) / piv;
                at(i, k) = l;
                if (l == 0.0) continue;
                int jend = std::min(n - 1, k + bw);
                for (int j = k + 1; j <= jend; ++j) at(i, j) -= l * at(k, j);
            }
        }
        ok = true;
    }
    void solve(const double* b, double* x) {
        std::vector<double> y(b, b + n);
        for (int i = 0; i < n; ++i) {
            int k0 = std::max(0, i - bw);
            double s = y[i];
            for (int k = k0; k < i; ++k) s -= at(i, k) * y[k];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            int k1 = std::min(n - 1, i + bw);
            double s = y[i];
            for (int k = i + 1; k <= k1; ++k) s -= at(i, k) * x[k];
            x[i] = s / at(i, i);
        }
    }
};

// The Newton-Raphson loop of OdeSolverEulerImplicit::solve (Math/OdeSolverEulerImplicit.cpp:55-122), written
// once over the equation: assemble(xn, vn) forms the system matrix and right-hand side of assembleLinearSystem
// (:124-161) at the estimate (xn, vn), solve(sol) solves that system. newState starts as the current state
// (xn = xs, vn = vs on entry); each iteration adds the solution to vn and sets xn = xs + dt vn; with more than one
// iteration allowed, the loop stops once ||sol - sol_prev||_inf < eps, sol_prev starting at 0. Returns the
// number of linear solves. Sim::solve_implicit and the harness of the reference's non-linear ODE test both use
// it.
template <class Assemble, class Solve>
static long newton_implicit_euler(double dt, long max_iter, double eps, const std::vector<double>& xs,
                                  std::vector<double>& xn, std::vector<double>& vn, std::vector<double>& sol,
                                  Assemble&& assemble, Solve&& solve) {
    std::vector<double> prev(xs.size(), 0.0);
    long it = 0, solves = 0;
    while (it < max_iter) {
        assemble(xn, vn);
        solve(sol);
        ++solves;
        for (size_t i = 0; i < vn.size(); ++i) vn[i] += sol[i];
        for (size_t i = 0; i < xn.size(); ++i) xn[i] = xs[i] + dt * vn[i];
        if (max_iter > 1) {
            double var = 0.0;
            for (size_t i = 0; i < sol.size(); ++i) var = std::max(var, std::fabs(sol[i] - prev[i]));
            if (var < eps) break;
            prev = sol;
        }
        it++;
    }
    return solves;
}

// Math/OdeSolver.h:36-49 (numbered as the GPU library numbers them: the two implicit schemes first)
enum Scheme {
    EULER_IMPLICIT = 0,
    LINEAR_EULER_IMPLICIT = 1,
    STATIC = 2,
    LINEAR_STATIC = 3,
    EULER_EXPLICIT = 4,
    LINEAR_EULER_EXPLICIT = 5,
    EULER_EXPLICIT_MODIFIED = 6,
    LINEAR_EULER_EXPLICIT_MODIFIED = 7,
    RUNGE_KUTTA_4 = 8,
    LINEAR_RUNGE_KUTTA_4 = 9
};
// OdeEquation update flags (Math/OdeEquation.h:34-41)
enum { UPD_F = 1, UPD_M = 2, UPD_D = 4, UPD_K = 8, UPD_FMDK = 15 };
enum LinSolver { LU = 0, CG = 1 };

// ------------------------------------------------------------------------------------------
// Representation + ODE solver (Physics/FemRepresentation.cpp, Physics/DeformableRepresentation.cpp,
// Math/OdeSolverEulerImplicit.cpp, Math/OdeSolverLinearEulerImplicit.cpp, Math/OdeState.cpp).
// ------------------------------------------------------------------------------------------
struct Sim {
    int num_nodes = 0;
    std::vector<Element> elems;
    std::vector<double> x0, x, v;          // current state
    std::vector<int> bc_dofs;              // OdeState::m_boundaryConditionsAsDofIds (insertion order)
    std::vector<char> bc_mask;
    std::vector<double> mass_per_node;
    double gravity[3] = {0.0, -9.81, 0.0};  // Physics/Representation.cpp:36-38
    bool gravity_enabled = true;
    std::vector<double> ext_force;          // external generalized force (constant while set)
    bool has_ext = false;
    // Rayleigh damping and mass lumping (Physics/FemRepresentation.h:89-142)
    double ray_mass = 0.0, ray_stiff = 0.0;
    bool lumping = false;
    int init_state = 0;                     // OdeEquation::m_initState (Math/OdeEquation.cpp:71-104)
    // Localized external generalized force, stiffness and damping of the next update
    // (Physics/Fem3DRepresentation.cpp:133-198), reset after it (Physics/DeformableRepresentation.cpp:236-253).
    // Stiffness and damping are kept per node block (row node, column node), 3x3 row-major.
    std::vector<double> loc_f;
    std::map<std::pair<int, int>, std::array<double, 9>> loc_K, loc_D;
    bool has_loc = false;
    Csc M, D, K, A;
    std::vector<double> f, rhs, solution;
    int scheme = EULER_IMPLICIT;
    int linsolver = CG;
    double cg_tol = std::numeric_limits<double>::epsilon();  // Eigen IterativeSolverBase default
    long cg_max_iter = -1;                                    // -> 2n
    long newton_max_iter = 1;                                 // Math/OdeSolverEulerImplicit.cpp:26
    double newton_eps = 1e-5;
    bool faithful_bc = false;   // O(#BC * n) zeroRow/zeroColumn as in Math/SparseMatrix.h:352-377
    bool linear_initialized = false;
    // OdeSolverLinearEulerImplicit forms the dense compliance on its first step (solve(..., true),
    // Math/OdeSolverLinearEulerImplicit.cpp:49); off unless asked for (it is D x D)
    bool compliance_at_first_linear_step = false;
    std::vector<double> compliance;  // D x D, column-major, of the last system matrix
    Csc Kconst;
    BandLU lu;
    Counters cnt;
    CgResult last_cg;
    long total_cg_iters = 0;
    long last_newton_iters = 0;  // linear solves of the last implicit step's Newton-Raphson loop
    bool valid = true;

    void init() {
        // FemRepresentation::doInitialize, Physics/FemRepresentation.cpp:85-132
        for (auto& e : elems) element_init(e, x0.data(), &cnt);
        mass_per_node.assign(num_nodes, 0.0);
        for (const auto& e : elems) {
            double mass = e.volume * e.rho;  // FemElement::getMass (FemElement.cpp:106-109)
            for (int a = 0; a < e.nn; ++a) mass_per_node[e.ids[a]] += mass / e.nn;
        }
        build_pattern(num_nodes, elems, &M);
        D = M;
        K = M;
        A = M;
        x = x0;
        v.assign(x0.size(), 0.0);
        ext_force.assign(x0.size(), 0.0);
        loc_f.assign(x0.size(), 0.0);
    }

    // Fem3DRepresentation::addExternalGeneralizedForce (Physics/Fem3DRepresentation.cpp:133-198):
    // force * coordinate[a] on node a; coordinate[a] * coordinate[b] * K (and D) on node block (a, b).
    void add_localized_force(int element, const double* coord, const double* force, const double* Kl,
                             const double* Dl) {
        if (element < 0 || element >= (int)elems.size()) throw std::runtime_error("Invalid localization");
        const Element& e = elems[element];
        for (int a = 0; a < e.nn; ++a)
            for (int c = 0; c < 3; ++c) loc_f[3 * (size_t)e.ids[a] + c] += force[c] * coord[a];
        if (Kl || Dl) {
            for (int a = 0; a < e.nn; ++a)
                for (int b = 0; b < e.nn; ++b) {
                    const auto key = std::make_pair(e.ids[a], e.ids[b]);
                    const double w = coord[a] * coord[b];
                    if (Kl) {
                        auto it = loc_K.find(key);
                        if (it == loc_K.end()) it = loc_K.emplace(key, std::array<double, 9>{}).first;
                        for (int k = 0; k < 9; ++k) it->second[k] += w * Kl[k];
                    }
                    if (Dl) {
                        auto it = loc_D.find(key);
                        if (it == loc_D.end()) it = loc_D.emplace(key, std::array<double, 9>{}).first;
                        for (int k = 0; k < 9; ++k) it->second[k] += w * Dl[k];
                    }
                }
        }
        has_loc = true;
    }
    // DeformableRepresentation::afterUpdate (Physics/DeformableRepresentation.cpp:236-253)
    void clear_localized() {
        std::fill(loc_f.begin(), loc_f.end(), 0.0);
        loc_K.clear();
        loc_D.clear();
        has_loc = false;
    }
    // m_K += m_externalGeneralizedStiffness (and D): the blocks lie inside the element pattern.
    static void add_blocks(const std::map<std::pair<int, int>, std::array<double, 9>>& m, Csc* S) {
        for (const auto& kv : m) {
            const int I = kv.first.first, J = kv.first.second;
            for (int c = 0; c < 3; ++c) {
                const int idx = S->find(3 * I, 3 * J + c);
                if (idx < 0) throw std::runtime_error("external block outside the pattern");
                for (int r = 0; r < 3; ++r) S->val[idx + r] += kv.second[3 * r + c];
            }
        }
    }
    void add_external_force(std::vector<double>& fv) {
        if (has_ext)
            for (size_t i = 0; i < fv.size(); ++i) fv[i] += ext_force[i];
        if (has_loc)
            for (size_t i = 0; i < fv.size(); ++i) fv[i] += loc_f[i];
    }

    // FemRepresentation::addRayleighDampingForce (Physics/FemRepresentation.cpp:565-629), scale 1:
    // F -= alpha_M M v and F -= alpha_K K v with the global matrices, or element by element through
    // FemElement::addMatVec (Physics/FemElement.cpp:168-210).
    void rayleigh_force(std::vector<double>& fv, const std::vector<double>& vs, bool global_k, bool global_m) {
        const size_t n = vs.size();
        std::vector<double> tmp(n);
        auto element_pass = [&](double alpha, bool mass) {
            for (const auto& e : elems) {
                const int ne = 3 * e.nn;
                const std::vector<double>& Me = mass ? e.M : e.K;
                for (int i = 0; i < ne; ++i) {
                    double acc = 0.0;
                    for (int k = 0; k < ne; ++k) acc += Me[i * ne + k] * vs[3 * (size_t)e.ids[k / 3] + k % 3];
                    fv[3 * (size_t)e.ids[i / 3] + i % 3] += alpha * acc;
                }
            }
        };
        if (ray_mass != 0.0) {
            if (global_m) {
                if (lumping) {
                    for (size_t i = 0; i < n; ++i) fv[i] -= ray_mass * (M.val[M.find((int)i, (int)i)] * vs[i]);
                } else {
                    csc_matvec(M, vs.data(), tmp.data());
                    for (size_t i = 0; i < n; ++i) fv[i] -= ray_mass * tmp[i];
                }
            } else {
                element_pass(-ray_mass, true);
            }
        }
        if (ray_stiff != 0.0) {
            if (global_k) {
                csc_matvec(K, vs.data(), tmp.data());
                for (size_t i = 0; i < n; ++i) fv[i] -= ray_stiff * tmp[i];
            } else {
                element_pass(-ray_stiff, false);
            }
        }
    }

    // Mass lumping of computeFMDK (Physics/FemRepresentation.cpp:499-508): each row's sum (columns in
    // ascending order) replaces the row, on the diagonal; the pattern is kept.
    void lump_mass() {
        std::vector<double> rs(M.n, 0.0);
        for (int j = 0; j < M.n; ++j)
            for (int p = M.outer[j]; p < M.outer[j + 1]; ++p) rs[M.inner[p]] += M.val[p];
        std::fill(M.val.begin(), M.val.end(), 0.0);
        for (int i = 0; i < M.n; ++i) M.val[M.find(i, i)] = rs[i];
    }

    void add_bc_dof(int dof) {
        if (bc_mask.empty()) bc_mask.assign(3 * (size_t)num_nodes, 0);
        if (!bc_mask[dof]) {
            bc_mask[dof] = 1;
            bc_dofs.push_back(dof);
        }
    }

    // FemRepresentation::addGravityForce (Physics/FemRepresentation.cpp:641-662)
    void add_gravity(std::vector<double>& fv) {
        if (!gravity_enabled) return;
        for (int n = 0; n < num_nodes; ++n)
            for (int c = 0; c < 3; ++c) fv[3 * n + c] += gravity[c] * (1.0 * mass_per_node[n]);
    }

    // FemRepresentation::updateFMDK + computeFMDK (Physics/FemRepresentation.cpp:479-563); vs is the
    // velocity of the state the update is taken at (Rayleigh damping force).
    void update_fmdk(const std::vector<double>& xs, const std::vector<double>& vs) {
        for (auto& e : elems) element_update(e, xs.data(), UPD_FMDK, &cnt);
        f.assign(xs.size(), 0.0);
        std::fill(M.val.begin(), M.val.end(), 0.0);
        std::fill(D.val.begin(), D.val.end(), 0.0);
        std::fill(K.val.begin(), K.val.end(), 0.0);
        for (const auto& e : elems) {
            for (int a = 0; a < e.nn; ++a)
                for (int c = 0; c < 3; ++c) f[3 * e.ids[a] + c] += e.f[3 * a + c];
            add_element_matrix(e, e.M, 1.0, &M);
            // addDamping is a no-op for 3D elements (m_useDamping false, FemElement.cpp:131-145)
            add_element_matrix(e, e.K, 1.0, &K);
        }
        if (lumping) lump_mass();
        if (ray_mass != 0.0)
            for (size_t p = 0; p < D.val.size(); ++p) D.val[p] += M.val[p] * ray_mass;
        if (ray_stiff != 0.0)
            for (size_t p = 0; p < D.val.size(); ++p) D.val[p] += K.val[p] * ray_stiff;
        add_gravity(f);
        rayleigh_force(f, vs, (init_state & UPD_K) != 0, true);
        add_external_force(f);
        if (has_loc) {
            add_blocks(loc_K, &K);
            add_blocks(loc_D, &D);
        }
        init_state |= UPD_FMDK;
    }

    // FemRepresentation::computeF (Physics/FemRepresentation.cpp:396-410): gravity, Rayleigh damping
    // force, element forces, external force. The elements were updated by the caller.
    void compute_f(const std::vector<double>& vs) {
        f.assign(vs.size(), 0.0);
        add_gravity(f);
        rayleigh_force(f, vs, (init_state & UPD_K) != 0, (init_state & UPD_F) != 0);
        for (const auto& e : elems)
            for (int a = 0; a < e.nn; ++a)
                for (int c = 0; c < 3; ++c) f[3 * e.ids[a] + c] += e.f[3 * a + c];
        add_external_force(f);
    }
    // FemRepresentation::computeM (:412-421): element masses, never lumped.
    void compute_m() {
        std::fill(M.val.begin(), M.val.end(), 0.0);
        for (const auto& e : elems) add_element_matrix(e, e.M, 1.0, &M);
    }
    // FemRepresentation::computeK (:460-477)
    void compute_k() {
        std::fill(K.val.begin(), K.val.end(), 0.0);
        for (const auto& e : elems) add_element_matrix(e, e.K, 1.0, &K);
        if (has_loc) add_blocks(loc_K, &K);
    }
    // OdeEquation::updateFMDK with options other than FMDK (Math/OdeEquation.cpp:78-103), after
    // FemRepresentation::updateFMDK's element loop (Physics/FemRepresentation.cpp:547-563).
    void update_parts(const std::vector<double>& xs, const std::vector<double>& vs, int options) {
        for (auto& e : elems) element_update(e, xs.data(), options, &cnt);
        if (options & UPD_F) {
            compute_f(vs);
            init_state |= UPD_F;
        }
        if (options & UPD_M) {
            compute_m();
            init_state |= UPD_M;
        }
        if (options & UPD_K) {
            compute_k();
            init_state |= UPD_K;
        }
    }

    // OdeState::applyBoundaryConditionsToMatrix(SparseMatrix*) (Math/OdeState.cpp:197-218)
    void apply_bc_matrix(Csc* S) {
        if (faithful_bc) {
            // zeroRow / zeroColumn as written: every coefficient of the row (column) is looked up.
            for (int dof : bc_dofs) {
                for (int c = 0; c < S->n; ++c) {
                    int idx = S->find(dof, c);
                    if (idx >= 0 && S->val[idx] != 0.0) S->val[idx] = 0.0;
                }
                for (int r = 0; r < S->n; ++r) {
                    int idx = S->find(r, dof);
                    if (idx >= 0 && S->val[idx] != 0.0) S->val[idx] = 0.0;
                }
                S->val[S->find(dof, dof)] = 1.0;
            }
            return;
        }
        if (bc_dofs.empty()) return;
        for (int j = 0; j < S->n; ++j)
            for (int p = S->outer[j]; p < S->outer[j + 1]; ++p)
                if (bc_mask[j] || bc_mask[S->inner[p]]) S->val[p] = 0.0;
        for (int dof : bc_dofs) S->val[S->find(dof, dof)] = 1.0;
    }
    void apply_bc_vector(std::vector<double>& b) {
        for (int dof : bc_dofs) b[dof] = 0.0;
    }

    // OdeSolverEulerImplicit::assembleLinearSystem (Math/OdeSolverEulerImplicit.cpp:124-161)
    void assemble(double dt, const std::vector<double>& xs, const std::vector<double>& vs,
                  const std::vector<double>& xn, const std::vector<double>& vn) {
        update_fmdk(xn, vn);
        double idt = 1.0 / dt;
        for (size_t p = 0; p < A.val.size(); ++p) A.val[p] = M.val[p] * idt + D.val[p] + K.val[p] * dt;
        if (bc_mask.empty()) bc_mask.assign(xs.size(), 0);
        apply_bc_matrix(&A);
        if (linsolver == LU) lu.factor(A);
        size_t n = xs.size();
        std::vector<double> t(n), Kt(n), Mt(n);
        for (size_t i = 0; i < n; ++i) t[i] = (xn[i] - xs[i]) - vn[i] * dt;
        csc_matvec(K, t.data(), Kt.data());
        rhs.assign(n, 0.0);
        for (size_t i = 0; i < n; ++i) rhs[i] = f[i] + Kt[i];
        for (size_t i = 0; i < n; ++i) t[i] = vn[i] - vs[i];
        csc_matvec(M, t.data(), Mt.data());
        for (size_t i = 0; i < n; ++i) rhs[i] -= Mt[i] / dt;
        apply_bc_vector(rhs);
    }

    void linear_solve(const std::vector<double>& b, std::vector<double>& xsol) {
        xsol.assign(b.size(), 0.0);
        if (linsolver == LU) {
            lu.solve(b.data(), xsol.data());
            last_cg = CgResult();
        } else {
            last_cg = cg_solve_lower(A, b.data(), xsol.data(), cg_tol, cg_max_iter);
            total_cg_iters += last_cg.iters;
        }
    }

    // OdeSolver::computeComplianceMatrixFromSystemMatrix (Math/OdeSolver.cpp:69-76): the inverse of the last
    // system matrix (LinearSparseSolveAndInverseLU::getInverse solves the identity through the factorization,
    // Math/LinearSparseSolveAndInverse.cpp:86-90; the CG solver's getInverse is the dense inverse, :25-29, here
    // through the same LU), then boundary rows and columns zeroed with a zero diagonal
    // (OdeState::applyBoundaryConditionsToMatrix(&C, false), Math/OdeState.cpp:175-195).
    void compute_compliance() {
        const int n = A.n;
        if (linsolver != LU) lu.factor(A);
        compliance.assign((size_t)n * n, 0.0);
        std::vector<double> e(n, 0.0);
        for (int j = 0; j < n; ++j) {
            e[j] = 1.0;
            lu.solve(e.data(), compliance.data() + (size_t)j * n);
            e[j] = 0.0;
        }
        for (int dof : bc_dofs) {
            for (int j = 0; j < n; ++j) compliance[(size_t)j * n + dof] = 0.0;
            for (int i = 0; i < n; ++i) compliance[(size_t)dof * n + i] = 0.0;
        }
    }

    // OdeSolverEulerImplicit::solve (Math/OdeSolverEulerImplicit.cpp:55-122), no compliance.
    void solve_implicit(double dt) {
        std::vector<double> xs = x, vs = v;
        std::vector<double> xn = x, vn = v;
        last_newton_iters = newton_implicit_euler(
            dt, newton_max_iter, newton_eps, xs, xn, vn, solution,
            [&](const std::vector<double>& xe, const std::vector<double>& ve) { assemble(dt, xs, vs, xe, ve); },
            [&](std::vector<double>& sol) { linear_solve(rhs, sol); });
        x = xn;
        v = vn;
    }

    // OdeSolverLinearEulerImplicit::solve (Math/OdeSolverLinearEulerImplicit.cpp:43-64)
    // with DeformableRepresentation::applyCompliance (Physics/DeformableRepresentation.cpp:175-190).
    void solve_linear_implicit(double dt) {
        if (!linear_initialized) {
            solve_implicit(dt);
            Kconst = K;  // pruned(): drops exact zeros only, which leaves every product unchanged
            linear_initialized = true;
            if (compliance_at_first_linear_step) compute_compliance();
            return;
        }
        update_parts(x, v, UPD_F);
        size_t n = x.size();
        std::vector<double> vdt(n), Kv(n);
        for (size_t i = 0; i < n; ++i) vdt[i] = v[i] * dt;
        csc_matvec(Kconst, vdt.data(), Kv.data());
        std::vector<double> b(n);
        for (size_t i = 0; i < n; ++i) b[i] = f[i] - Kv[i];
        apply_bc_vector(b);
        std::vector<double> dv;
        linear_solve(b, dv);
        apply_bc_vector(dv);
        for (size_t i = 0; i < n; ++i) v[i] = v[i] + dv[i];
        for (size_t i = 0; i < n; ++i) x[i] = x[i] + dt * v[i];
    }

    // ---- integrators other than implicit Euler (Math/OdeSolver*.cpp) ----
    // System matrix M / dt (explicit schemes, RK4) or K (static), boundary conditions applied and
    // handed to the linear solver (m_linearSolver->setMatrix).
    void set_system(bool stiffness, double dt) {
        if (stiffness) A.val = K.val;
        else
            for (size_t p = 0; p < A.val.size(); ++p) A.val[p] = M.val[p] / dt;
        if (bc_mask.empty()) bc_mask.assign(x.size(), 0);
        apply_bc_matrix(&A);
        if (linsolver == LU) lu.factor(A);
    }
    std::vector<double> rhs_from_f(double div = 1.0) {
        std::vector<double> b = f;
        apply_bc_vector(b);
        if (div != 1.0)
            for (double& q : b) q /= div;
        return b;
    }
    // DeformableRepresentation::applyCompliance (Physics/DeformableRepresentation.cpp:175-190)
    std::vector<double> apply_compliance(std::vector<double> b) {
        apply_bc_vector(b);
        std::vector<double> out;
        linear_solve(b, out);
        apply_bc_vector(out);
        return out;
    }

    // OdeSolverStatic::solve / assembleLinearSystem (Math/OdeSolverStatic.cpp:32-70); the linear variant
    // (Math/OdeSolverLinearStatic.cpp:33-50) keeps the first system and applies its compliance to F.
    void solve_static(double dt, bool linear) {
        std::vector<double> dx;
        if (linear && linear_initialized) {
            update_parts(x, v, UPD_F);
            dx = apply_compliance(f);
        } else {
            update_parts(x, v, UPD_F | UPD_K);
            set_system(true, dt);
            rhs = rhs_from_f();
            linear_solve(rhs, dx);
            solution = dx;
            linear_initialized = linear;
        }
        for (size_t i = 0; i < x.size(); ++i) x[i] = x[i] + dx[i];
        std::fill(v.begin(), v.end(), 0.0);
    }

    // OdeSolverEulerExplicit (Math/OdeSolverEulerExplicit.cpp:32-70) and OdeSolverEulerExplicitModified
    // (Math/OdeSolverEulerExplicitModified.cpp:32-71): (M / dt) dv = F; the explicit scheme moves x with
    // the old velocity, the modified one with the new. Linear variants (Math/OdeSolverLinearEulerExplicit.cpp,
    // Math/OdeSolverLinearEulerExplicitModified.cpp): later steps apply the first step's compliance.
    void solve_explicit(double dt, bool modified, bool linear) {
        std::vector<double> dv;
        if (linear && linear_initialized) {
            update_parts(x, v, UPD_F);
            dv = apply_compliance(f);
        } else {
            update_parts(x, v, UPD_F | UPD_M);
            set_system(false, dt);
            rhs = rhs_from_f();
            linear_solve(rhs, dv);
            solution = dv;
            linear_initialized = linear;
        }
        const size_t n = x.size();
        if (modified) {
            for (size_t i = 0; i < n; ++i) v[i] = v[i] + dv[i];
            for (size_t i = 0; i < n; ++i) x[i] = x[i] + dt * v[i];
        } else {
            for (size_t i = 0; i < n; ++i) x[i] = x[i] + dt * v[i];
            for (size_t i = 0; i < n; ++i) v[i] = v[i] + dv[i];
        }
    }

    // OdeSolverRungeKutta4::solve (Math/OdeSolverRungeKutta4.cpp:32-112); the linear variant
    // (Math/OdeSolverLinearRungeKutta4.cpp:33-91) evaluates every stage through the first step's compliance.
    void solve_rk4(double dt, bool linear) {
        const size_t n = x.size();
        const bool first = !(linear && linear_initialized);
        std::vector<double> xn(n), vn(n), k1v = v, k2v, k3v, k4v, k1a, k2a, k3a, k4a;
        if (first) {
            update_parts(x, v, UPD_F | UPD_M);
            set_system(false, dt);
            rhs = rhs_from_f();
            std::vector<double> b = rhs;
            for (double& q : b) q /= dt;
            linear_solve(b, k1a);
        } else {
            update_parts(x, v, UPD_F);
            std::vector<double> b = f;
            for (double& q : b) q /= dt;
            k1a = apply_compliance(b);
        }
        auto stage = [&](const std::vector<double>& kv, const std::vector<double>& ka, double h,
                         std::vector<double>& vout, std::vector<double>& aout) {
            for (size_t i = 0; i < n; ++i) xn[i] = x[i] + kv[i] * h;
            for (size_t i = 0; i < n; ++i) vn[i] = v[i] + ka[i] * h;
            vout = vn;
            update_parts(xn, vn, UPD_F);
            if (first) {
                std::vector<double> b = rhs_from_f(dt);
                linear_solve(b, aout);
            } else {
                std::vector<double> b = f;
                for (double& q : b) q /= dt;
                aout = apply_compliance(b);
            }
        };
        // y + k dt / 2: Eigen evaluates (k * dt) / 2.0 per coefficient
        auto half = [&](const std::vector<double>& kv, const std::vector<double>& ka, std::vector<double>& vout,
                        std::vector<double>& aout) {
            for (size_t i = 0; i < n; ++i) xn[i] = x[i] + kv[i] * dt / 2.0;
            for (size_t i = 0; i < n; ++i) vn[i] = v[i] + ka[i] * dt / 2.0;
            vout = vn;
            update_parts(xn, vn, UPD_F);
            if (first) {
                std::vector<double> b = rhs_from_f(dt);
                linear_solve(b, aout);
            } else {
                std::vector<double> b = f;
                for (double& q : b) q /= dt;
                aout = apply_compliance(b);
            }
        };
        half(k1v, k1a, k2v, k2a);
        half(k2v, k2a, k3v, k3a);
        stage(k3v, k3a, dt, k4v, k4a);
        for (size_t i = 0; i < n; ++i) x[i] = x[i] + (k1v[i] + k4v[i] + 2.0 * (k2v[i] + k3v[i])) * dt / 6.0;
        for (size_t i = 0; i < n; ++i) v[i] = v[i] + (k1a[i] + k4a[i] + 2.0 * (k2a[i] + k3a[i])) * dt / 6.0;
        if (first) linear_initialized = linear;
    }

    // ---- adjoint (reference gradient for the GPU adjoint; no counterpart in the reference) ----
    // Element K, f at local positions xe through the literal element code (a copy of the element, so
    // the simulation's own element state is untouched).
    void element_kf(const Element& e0, const double* xe, std::vector<double>& Ke, std::vector<double>& fe) const {
        Element e = e0;
        Counters tmp;
        const int n = 3 * e.nn;
        if (e.kind == TET_COROTATIONAL) {
            coro_rotation_mass_stiffness(e, xe, true, false, true, &tmp);
            double Rx0[12];
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < 3; ++i)
                    Rx0[3 * a + i] = e.R.m[i][0] * e.x0[3 * a] + e.R.m[i][1] * e.x0[3 * a + 1] + e.R.m[i][2] * e.x0[3 * a + 2];
            Ke = e.K;
            fe.assign(n, 0.0);
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += e.K[i * n + k] * (xe[k] - Rx0[k]);
                fe[i] = -acc;
            }
        } else {
            Ke = e.K;
            fe.assign(n, 0.0);
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += e.K[i * n + k] * (xe[k] - e.x0[k]);
                fe[i] = -acc;
            }
        }
    }

    // phi_e(xe) = mu.(f - dt K v) - dt mu^T S_K dv, S_K the lower-triangle self-adjoint view by global DOF
    // index of the element's K (entry (r, c) = K(r, c) if r >= c else K(c, r)).
    double element_phi(const Element& e, const double* xe, const double* ve, const double* me, const double* dve,
                       double dt, std::vector<double>* Kout = nullptr) const {
        std::vector<double> Ke, fe;
        element_kf(e, xe, Ke, fe);
        const int n = 3 * e.nn;
        double phi = 0.0;
        for (int i = 0; i < n; ++i) phi += me[i] * fe[i];
        for (int i = 0; i < n; ++i) {
            const int gi = 3 * e.ids[i / 3] + i % 3;
            for (int j = 0; j < n; ++j) {
                const int gj = 3 * e.ids[j / 3] + j % 3;
                phi -= dt * me[i] * Ke[i * n + j] * ve[j];
                const double sij = gi >= gj ? Ke[i * n + j] : Ke[j * n + i];
                phi -= dt * me[i] * sij * dve[j];
            }
        }
        if (Kout) *Kout = Ke;
        return phi;
    }

    // Reverse of one implicit-Euler step (one Newton iteration; CG on the lower-triangle view) from
    // (xn, vn) to vnext: lx, lv = dL/dx', dL/dv' -> lx_out, lv_out = dL/dx, dL/dv; gradE += dL/dE.
    //   gv = lv + dt lx;  S mu = gv (dense solve);  mu~ = mu with BC entries zero;
    //   lx_out = lx + d phi/dx;  lv_out = gv - dt K^T mu~;  dL/dE_e += phi_e / E_e
    // with d phi_e / d x_e by Richardson-extrapolated central differences of element_phi.
    void adjoint_step(double dt, const double* xn, const double* vn, const double* vnext, const double* lx,
                      const double* lv, double* lx_out, double* lv_out, double* gradE) {
        const size_t n = x0.size();
        std::vector<double> xs(xn, xn + n), vs(vn, vn + n);
        assemble(dt, xs, vs, xs, vs);
        std::vector<double> gv(n), mu(n, 0.0);
        for (size_t i = 0; i < n; ++i) gv[i] = lv[i] + dt * lx[i];
        // dense lower-triangle self-adjoint view of A and Gaussian elimination with partial pivoting
        std::vector<double> S(n * n, 0.0);
        for (size_t j = 0; j < n; ++j)
            for (int p = A.outer[j]; p < A.outer[j + 1]; ++p) {
                const size_t i = (size_t)A.inner[p];
                if (i >= j) {
                    S[i * n + j] = A.val[p];
                    S[j * n + i] = A.val[p];
                }
            }
        std::vector<double> rhsv = gv;
        std::vector<size_t> perm(n);
        for (size_t k = 0; k < n; ++k) {
            size_t piv = k;
            for (size_t i = k + 1; i < n; ++i)
                if (std::fabs(S[i * n + k]) > std::fabs(S[piv * n + k])) piv = i;
            if (piv != k) {
                for (size_t j = 0; j < n; ++j) std::swap(S[k * n + j], S[piv * n + j]);
                std::swap(rhsv[k], rhsv[piv]);
            }
            for (size_t i = k + 1; i < n; ++i) {
                const double fct = S[i * n + k] / S[k * n + k];
                if (fct == 0.0) continue;
                for (size_t j = k; j < n; ++j) S[i * n + j] -= fct * S[k * n + j];
                rhsv[i] -= fct * rhsv[k];
            }
        }
        for (size_t k = n; k-- > 0;) {
            double acc = rhsv[k];
            for (size_t j = k + 1; j < n; ++j) acc -= S[k * n + j] * mu[j];
            mu[k] = acc / S[k * n + k];
        }
        for (int d : bc_dofs) mu[d] = 0.0;
        std::vector<double> gx_n(n, 0.0), gv_n(n, 0.0);
        for (size_t ei = 0; ei < elems.size(); ++ei) {
            const Element& e = elems[ei];
            const int ne = 3 * e.nn;
            double xe[24], ve[24], me[24], dve[24];
            double ext = 0.0;
            for (int a = 0; a < e.nn; ++a)
                for (int c = 0; c < 3; ++c) {
                    const size_t d = 3 * (size_t)e.ids[a] + c;
                    xe[3 * a + c] = xn[d];
                    ve[3 * a + c] = vn[d];
                    me[3 * a + c] = mu[d];
                    dve[3 * a + c] = vnext[d] - vn[d];
                    ext = std::max(ext, std::fabs(e.x0[3 * a + c] - e.x0[c]));
                }
            std::vector<double> Ke;
            const double phi = element_phi(e, xe, ve, me, dve, dt, &Ke);
            gradE[ei] += phi / e.E;
            const double h = 1e-4 * ext;
            for (int k = 0; k < ne; ++k) {
                auto central = [&](double step) {
                    double xp[24], xm[24];
                    std::memcpy(xp, xe, sizeof(double) * ne);
                    std::memcpy(xm, xe, sizeof(double) * ne);
                    xp[k] += step;
                    xm[k] -= step;
                    return (element_phi(e, xp, ve, me, dve, dt) - element_phi(e, xm, ve, me, dve, dt)) / (2 * step);
                };
                const double d1 = central(h), d2 = central(0.5 * h);
                gx_n[3 * (size_t)e.ids[k / 3] + k % 3] += (4.0 * d2 - d1) / 3.0;
            }
            for (int j = 0; j < ne; ++j) {
                double acc = 0.0;
                for (int i = 0; i < ne; ++i) acc += Ke[i * ne + j] * me[i];
                gv_n[3 * (size_t)e.ids[j / 3] + j % 3] -= dt * acc;
            }
        }
        for (size_t i = 0; i < n; ++i) {
            lx_out[i] = lx[i] + gx_n[i];
            lv_out[i] = gv[i] + gv_n[i];
        }
    }

    // New Young's moduli with the same Poisson ratios and densities (element re-initialisation).
    void set_young(const double* E) {
        for (size_t i = 0; i < elems.size(); ++i) {
            elems[i].E = E[i];
            element_init(elems[i], x0.data(), &cnt);
        }
        linear_initialized = false;
    }

    // FemRepresentation::update (Physics/FemRepresentation.cpp:232-294) minus compliance.
    int step(double dt) {
        if (!valid) return 1;
        switch (scheme) {
            case LINEAR_EULER_IMPLICIT: solve_linear_implicit(dt); break;
            case STATIC: solve_static(dt, false); break;
            case LINEAR_STATIC: solve_static(dt, true); break;
            case EULER_EXPLICIT: solve_explicit(dt, false, false); break;
            case LINEAR_EULER_EXPLICIT: solve_explicit(dt, false, true); break;
            case EULER_EXPLICIT_MODIFIED: solve_explicit(dt, true, false); break;
            case LINEAR_EULER_EXPLICIT_MODIFIED: solve_explicit(dt, true, true); break;
            case RUNGE_KUTTA_4: solve_rk4(dt, false); break;
            case LINEAR_RUNGE_KUTTA_4: solve_rk4(dt, true); break;
            default: solve_implicit(dt); break;
        }
        clear_localized();
        // OdeState::isValid (Math/OdeState.cpp:268-278)
        double sx = 0.0, sv = 0.0;
        for (double q : x) sx += q;
        for (double q : v) sv += q;
        if (!std::isfinite(sx) || !std::isfinite(sv)) valid = false;
        return valid ? 0 : 1;
    }
};

}  // namespace orc

// ------------------------------------------------------------------------------------------
// C API for the tests (ctypes).
// ------------------------------------------------------------------------------------------
using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Polar decomposition with Eigen's computeRotationScaling convention.
void orc_rotation_scaling(const double* F9, double* R9, double* S9) {
    M3 F, R, S;
    for (int i = 0; i < 9; ++i) F.m[i / 3][i % 3] = F9[i];
    rotation_scaling(F, &R, &S);
    for (int i = 0; i < 9; ++i) {
        R9[i] = R.m[i / 3][i % 3];
        S9[i] = S.m[i / 3][i % 3];
    }
}

int orc_tet_linear_km(const double* x0, double E, double nu, double rho, double* K, double* M, double* vol) {
    try {
        check_material(E, nu, rho);
        tet_linear_stiffness(x0, E, nu, K, vol);
        tet_linear_mass(*vol, rho, M);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

int orc_cube_km(const double* x0, double E, double nu, double rho, double* K, double* M, double* vol) {
    try {
        check_material(E, nu, rho);
        *vol = cube_volume(x0);
        cube_mass(x0, rho, M);
        cube_stiffness(x0, E, nu, K);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// One corotational element: initialize at x0, then updateFMDK(FMDK) at x (12 values).
// Returns the rotation, M, K (row-major 12x12) and f. flags: bit0 det R invalid, bit1 G singular.
int orc_coro_element(const double* x0, const double* x, double E, double nu, double rho, double* R9,
                     double* M, double* K, double* f, int* flags) {
    try {
        Element e;
        e.kind = TET_COROTATIONAL;
        e.nn = 4;
        for (int a = 0; a < 4; ++a) e.ids[a] = a;
        e.E = E;
        e.nu = nu;
        e.rho = rho;
        Counters c0;
        element_init(e, x0, &c0);
        Counters c;
        element_update(e, x, 1 | 2 | 4 | 8, &c);
        for (int i = 0; i < 9; ++i) R9[i] = e.R.m[i / 3][i % 3];
        std::memcpy(M, e.M.data(), 144 * sizeof(double));
        std::memcpy(K, e.K.data(), 144 * sizeof(double));
        std::memcpy(f, e.f.data(), 12 * sizeof(double));
        *flags = (c.det_r_invalid ? 1 : 0) | (c.g_singular ? 2 : 0);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// orc_coro_element with the polar factor R9 (row-major) supplied instead of computed.
int orc_coro_element_given_rotation(const double* x0, const double* x, double E, double nu, double rho,
                                    const double* R9, double* M, double* K, double* f, int* flags) {
    try {
        Element e;
        e.kind = TET_COROTATIONAL;
        e.nn = 4;
        for (int a = 0; a < 4; ++a) e.ids[a] = a;
        e.E = E;
        e.nu = nu;
        e.rho = rho;
        Counters c0;
        element_init(e, x0, &c0);
        M3 R;
        for (int i = 0; i < 9; ++i) R.m[i / 3][i % 3] = R9[i];
        Counters c;
        element_update(e, x, 1 | 2 | 4 | 8, &c, &R);
        std::memcpy(M, e.M.data(), 144 * sizeof(double));
        std::memcpy(K, e.K.data(), 144 * sizeof(double));
        std::memcpy(f, e.f.data(), 12 * sizeof(double));
        *flags = (c.det_r_invalid ? 1 : 0) | (c.g_singular ? 2 : 0);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// The reference's OdeComplexNonLinear test equation (Math/UnitTests/MockObject.h:267-338: a = x |v|^2 with
// M = I, f = |v|^2 x, D = 2 x v^T, K = |v|^2 I, no boundary conditions) stepped twice from its test's initial
// state (x = LinSpaced(1.4, 5.67), v = LinSpaced(-0.4, -0.3), dt = 1e-3;
// Math/UnitTests/OdeSolverEulerImplicitTests.cpp:139-220) with the reference's default direct solver, by
// OdeSolverEulerImplicit::solve (linear_variant 0) or OdeSolverLinearEulerImplicit::solve (1: the first step
// as the non-linear solver, then the first step's K and system matrix kept, Math/OdeSolverLinearEulerImplicit.cpp:43-64).
// out: x1, v1, x2, v2, f(x2, v2) (3 each); solves: linear solves of the two steps.
int orc_ode_complex_nonlinear(int linear_variant, long max_iter, double eps, double* out, long* solves) {
    try {
        const double dt = 1e-3;
        const int n = 3;
        auto force = [](const std::vector<double>& x, const std::vector<double>& v, std::vector<double>& f) {
            const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            for (int i = 0; i < 3; ++i) f[i] = v2 * x[i];
        };
        BandLU lu;
        std::vector<double> Kd(9), rhs(n);
        // assembleLinearSystem at (xn, vn) for the step from (xs, vs)
        auto assemble = [&](const std::vector<double>& xs, const std::vector<double>& vs, const std::vector<double>& xn,
                            const std::vector<double>& vn) {
            const double v2 = vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2];
            Csc A;
            A.n = n;
            A.outer = {0, 3, 6, 9};
            A.inner = {0, 1, 2, 0, 1, 2, 0, 1, 2};
            A.val.assign(9, 0.0);
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    const double M = i == j ? 1.0 : 0.0, D = 2.0 * xn[i] * vn[j], K = i == j ? v2 : 0.0;
                    Kd[3 * i + j] = K;
                    A.val[3 * j + i] = M * (1.0 / dt) + D + K * dt;
                }
            lu.factor(A);
            std::vector<double> f(n);
            force(xn, vn, f);
            for (int i = 0; i < n; ++i) {
                double Kt = 0.0;
                for (int k = 0; k < n; ++k) Kt += Kd[3 * i + k] * ((xn[k] - xs[k]) - vn[k] * dt);
                rhs[i] = f[i] + Kt - (vn[i] - vs[i]) / dt;
            }
        };
        std::vector<double> x0(n), v0(n);
        for (int i = 0; i < n; ++i) {
            x0[i] = 1.4 + (5.67 - 1.4) * i / 2.0;
            v0[i] = -0.4 + (-0.3 + 0.4) * i / 2.0;
        }
        std::vector<double> sol(n);
        auto nonlinear_step = [&](const std::vector<double>& xs, const std::vector<double>& vs, std::vector<double>& xn,
                                  std::vector<double>& vn) {
            xn = xs;
            vn = vs;
            return newton_implicit_euler(
                dt, max_iter, eps, xs, xn, vn, sol,
                [&](const std::vector<double>& xe, const std::vector<double>& ve) { assemble(xs, vs, xe, ve); },
                [&](std::vector<double>& s) { lu.solve(rhs.data(), s.data()); });
        };
        std::vector<double> x1, v1, x2, v2;
        long ns = nonlinear_step(x0, v0, x1, v1);
        if (!linear_variant) {
            ns += nonlinear_step(x1, v1, x2, v2);
        } else {
            // f(x1, v1) - K_const (v1 dt), solved against the first step's last system matrix
            std::vector<double> f(n), b(n), dv(n);
            force(x1, v1, f);
            for (int i = 0; i < n; ++i) {
                double Kv = 0.0;
                for (int k = 0; k < n; ++k) Kv += Kd[3 * i + k] * (v1[k] * dt);
                b[i] = f[i] - Kv;
            }
            lu.solve(b.data(), dv.data());
            ns += 1;
            v2.resize(n);
            x2.resize(n);
            for (int i = 0; i < n; ++i) {
                v2[i] = v1[i] + dv[i];
                x2[i] = x1[i] + dt * v2[i];
            }
        }
        std::vector<double> f2(n);
        force(x2, v2, f2);
        for (int i = 0; i < n; ++i) {
            out[i] = x1[i];
            out[3 + i] = v1[i];
            out[6 + i] = x2[i];
            out[9 + i] = v2[i];
            out[12 + i] = f2[i];
        }
        *solves = ns;
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

void* orc_sim_create(int kind, int num_nodes, const double* x0, int num_elems, const int* conn, const double* E,
                     const double* nu, const double* rho) {
    try {
        auto* s = new Sim();
        s->num_nodes = num_nodes;
        s->x0.assign(x0, x0 + 3 * (size_t)num_nodes);
        int nn = (kind == HEX) ? 8 : 4;
        s->elems.resize(num_elems);
        for (int i = 0; i < num_elems; ++i) {
            Element& e = s->elems[i];
            e.kind = kind;
            e.nn = nn;
            for (int a = 0; a < nn; ++a) e.ids[a] = conn[(size_t)i * nn + a];
            e.E = E[i];
            e.nu = nu[i];
            e.rho = rho[i];
        }
        s->init();
        s->bc_mask.assign(3 * (size_t)num_nodes, 0);
        return s;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void orc_sim_destroy(void* h) { delete static_cast<Sim*>(h); }

void orc_sim_add_bc_dof(void* h, int dof) { static_cast<Sim*>(h)->add_bc_dof(dof); }

void orc_sim_set_gravity(void* h, double gx, double gy, double gz, int enabled) {
    Sim* s = static_cast<Sim*>(h);
    s->gravity[0] = gx;
    s->gravity[1] = gy;
    s->gravity[2] = gz;
    s->gravity_enabled = enabled != 0;
}

void orc_sim_set_external_force(void* h, const double* fext) {
    Sim* s = static_cast<Sim*>(h);
    if (fext) {
        s->ext_force.assign(fext, fext + s->x0.size());
        s->has_ext = true;
    } else {
        s->has_ext = false;
    }
}

void orc_sim_set_solver(void* h, int scheme, int linsolver, double tol, long max_iter, long newton_max_iter,
                        int faithful_bc) {
    Sim* s = static_cast<Sim*>(h);
    s->scheme = scheme;
    s->linsolver = linsolver;
    s->cg_tol = tol;
    s->cg_max_iter = max_iter;
    s->newton_max_iter = newton_max_iter;
    s->faithful_bc = faithful_bc != 0;
}

void orc_sim_set_state(void* h, const double* x, const double* v) {
    Sim* s = static_cast<Sim*>(h);
    if (x) s->x.assign(x, x + s->x0.size());
    if (v) s->v.assign(v, v + s->x0.size());
}

void orc_sim_get_state(void* h, double* x, double* v) {
    Sim* s = static_cast<Sim*>(h);
    if (x) std::memcpy(x, s->x.data(), s->x.size() * sizeof(double));
    if (v) std::memcpy(v, s->v.data(), s->v.size() * sizeof(double));
}

int orc_sim_step(void* h, double dt, int nsteps) {
    Sim* s = static_cast<Sim*>(h);
    try {
        for (int i = 0; i < nsteps; ++i)
            if (s->step(dt)) return 1;
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

long orc_sim_nnz(void* h) { return (long)static_cast<Sim*>(h)->M.inner.size(); }

void orc_sim_pattern(void* h, int* outer, int* inner) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(outer, s->M.outer.data(), s->M.outer.size() * sizeof(int));
    std::memcpy(inner, s->M.inner.data(), s->M.inner.size() * sizeof(int));
}

// Last assembled system matrix A (after boundary conditions), values in pattern order.
void orc_sim_system_values(void* h, double* vals) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(vals, s->A.val.data(), s->A.val.size() * sizeof(double));
}
// Rayleigh damping coefficients and mass lumping (FemRepresentation::setRayleighDampingMass /
// setRayleighDampingStiffness / setMassLumping, Physics/FemRepresentation.cpp:212-220, 320-323).
void orc_sim_set_damping(void* h, double rayleigh_mass, double rayleigh_stiffness, int lumping) {
    Sim* s = static_cast<Sim*>(h);
    s->ray_mass = rayleigh_mass;
    s->ray_stiff = rayleigh_stiffness;
    s->lumping = lumping != 0;
}

// Fem3DRepresentation::addExternalGeneralizedForce (Physics/Fem3DRepresentation.cpp:133-198) for the next
// step; K and D (3x3 row-major) may be null.
int orc_sim_add_localized_force(void* h, int element, const double* coord, const double* force, const double* K,
                                const double* D) {
    try {
        static_cast<Sim*>(h)->add_localized_force(element, coord, force, K, D);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Rebuild f, M, D, K at the current state as computeFMDK does (no step): for checks of the assembled terms.
int orc_sim_update_fmdk(void* h) {
    Sim* s = static_cast<Sim*>(h);
    try {
        s->update_fmdk(s->x, s->v);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// DeformableRepresentation::applyCompliance (Physics/DeformableRepresentation.cpp:175-190) with the last
// system matrix, column by column (num_columns column-major columns of ndof values).
int orc_sim_apply_compliance(void* h, const double* b, int num_columns, double* out) {
    Sim* s = static_cast<Sim*>(h);
    try {
        const size_t n = s->x.size();
        for (int j = 0; j < num_columns; ++j) {
            std::vector<double> col(b + (size_t)j * n, b + (size_t)(j + 1) * n);
            std::vector<double> r = s->apply_compliance(col);
            std::memcpy(out + (size_t)j * n, r.data(), n * sizeof(double));
        }
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

void orc_sim_matrix_values(void* h, int which, double* vals) {
    Sim* s = static_cast<Sim*>(h);
    const Csc& S = which == 0 ? s->M : (which == 1 ? s->K : which == 3 ? s->D : s->A);
    std::memcpy(vals, S.val.data(), S.val.size() * sizeof(double));
}
void orc_sim_rhs(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(out, s->rhs.data(), s->rhs.size() * sizeof(double));
}
void orc_sim_force(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(out, s->f.data(), s->f.size() * sizeof(double));
}
void orc_sim_mass_per_node(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(out, s->mass_per_node.data(), s->mass_per_node.size() * sizeof(double));
}
long orc_sim_last_cg_iters(void* h) { return static_cast<Sim*>(h)->last_cg.iters; }
double orc_sim_last_cg_error(void* h) { return static_cast<Sim*>(h)->last_cg.error; }
long orc_sim_total_cg_iters(void* h) { return static_cast<Sim*>(h)->total_cg_iters; }
long orc_sim_last_newton_iters(void* h) { return static_cast<Sim*>(h)->last_newton_iters; }
void orc_sim_set_compliance_at_first_linear_step(void* h, int on) {
    static_cast<Sim*>(h)->compliance_at_first_linear_step = on != 0;
}
// The dense compliance of the last system matrix (computes it now); out: D x D column-major.
int orc_sim_compliance_matrix(void* h, double* out) {
    try {
        Sim* s = static_cast<Sim*>(h);
        if (s->A.n == 0) throw std::runtime_error("no system matrix yet");
        s->compute_compliance();
        std::memcpy(out, s->compliance.data(), s->compliance.size() * sizeof(double));
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}
// OdeSolverEulerImplicit::setNewtonRaphsonEpsilonConvergence (Math/OdeSolverEulerImplicit.cpp:43-48)
void orc_sim_set_newton_eps(void* h, double eps) { static_cast<Sim*>(h)->newton_eps = eps; }
void orc_sim_counters(void* h, long* det_r, long* g_sing) {
    Sim* s = static_cast<Sim*>(h);
    *det_r = s->cnt.det_r_invalid;
    *g_sing = s->cnt.g_singular;
}
int orc_sim_valid(void* h) { return static_cast<Sim*>(h)->valid ? 1 : 0; }

// Element data after the last update: R (9), K (nn3^2), f (nn3) of element i.
void orc_sim_element(void* h, int i, double* R9, double* K, double* f) {
    Sim* s = static_cast<Sim*>(h);
    const Element& e = s->elems[i];
    if (R9)
        for (int k = 0; k < 9; ++k) R9[k] = e.R.m[k / 3][k % 3];
    if (K) std::memcpy(K, e.K.data(), e.K.size() * sizeof(double));
    if (f) std::memcpy(f, e.f.data(), e.f.size() * sizeof(double));
}

int orc_sim_adjoint_step(void* h, double dt, const double* xn, const double* vn, const double* vnext,
                         const double* lx, const double* lv, double* lx_out, double* lv_out, double* gradE) {
    try {
        static_cast<Sim*>(h)->adjoint_step(dt, xn, vn, vnext, lx, lv, lx_out, lv_out, gradE);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

int orc_sim_set_young(void* h, const double* E) {
    try {
        static_cast<Sim*>(h)->set_young(E);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Standalone Eigen-CG (lower self-adjoint view) on a CSC matrix.
long orc_cg_solve(int n, const int* outer, const int* inner, const double* vals, const double* b, double* x,
                  double tol, long max_iter, double* err) {
    Csc A;
    A.n = n;
    A.outer.assign(outer, outer + n + 1);
    A.inner.assign(inner, inner + outer[n]);
    A.val.assign(vals, vals + outer[n]);
    CgResult r = cg_solve_lower(A, b, x, tol, max_iter);
    if (err) *err = r.error;
    return r.iters;
}

}  // extern "C"
