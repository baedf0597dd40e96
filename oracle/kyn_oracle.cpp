// ORACLE — test infrastructure, see kyn_oracle.hpp. CPU FP64 restatement of the reference
// beam path; every function names the reference file:line it follows.
#include "kyn_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>

namespace kyn_oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;

// C = A * B (3x3)
Mat3 mm(const Mat3& A, const Mat3& B) {
    Mat3 C{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
    return C;
}
// C = A^T * B
Mat3 mtm(const Mat3& A, const Mat3& B) {
    Mat3 C{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.;
            for (int k = 0; k < 3; ++k) s += A[k * 3 + i] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
    return C;
}
// C = A * B^T
Mat3 mmt(const Mat3& A, const Mat3& B) {
    Mat3 C{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[j * 3 + k];
            C[i * 3 + j] = s;
        }
    return C;
}
Vec3 mv(const Mat3& A, const Vec3& x) {
    Vec3 y{};
    for (int i = 0; i < 3; ++i) y[i] = A[i * 3] * x[0] + A[i * 3 + 1] * x[1] + A[i * 3 + 2] * x[2];
    return y;
}
Vec3 mtv(const Mat3& A, const Vec3& x) {
    Vec3 y{};
    for (int i = 0; i < 3; ++i) y[i] = A[i] * x[0] + A[3 + i] * x[1] + A[6 + i] * x[2];
    return y;
}
Mat3 sub3(const Mat6& M, int r, int c) {
    Mat3 S{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i * 3 + j] = M[(r + i) * 6 + c + j];
    return S;
}
void put3(Mat6& M, int r, int c, const Mat3& S) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[(r + i) * 6 + c + j] = S[i * 3 + j];
}
int count_active_dofs(int sig) {  // src/dof_management/freedom_signature.hpp:38-46
    int c = 0;
    for (int v = sig; v > 0; v >>= 1) c += v & 1;
    return c;
}
constexpr int kAll = 0b111111;
constexpr int kJustPosition = 0b111000;  // freedom_signature.hpp:20
constexpr int kNone = 0;
}  // namespace

// ---------------------------------------------------------------------------------------
// Math
// ---------------------------------------------------------------------------------------
// src/math/quaternion_operations.hpp:19-32
Mat3 QuaternionToRotationMatrix(const Quat& q) {
    Mat3 R{};
    R[0] = q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
    R[1] = 2. * (q[1] * q[2] - q[0] * q[3]);
    R[2] = 2. * (q[1] * q[3] + q[0] * q[2]);
    R[3] = 2. * (q[1] * q[2] + q[0] * q[3]);
    R[4] = q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3];
    R[5] = 2. * (q[2] * q[3] - q[0] * q[1]);
    R[6] = 2. * (q[1] * q[3] - q[0] * q[2]);
    R[7] = 2. * (q[2] * q[3] + q[0] * q[1]);
    R[8] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
    return R;
}
// src/math/quaternion_operations.hpp:113-125
Vec3 RotateVectorByQuaternion(const Quat& q, const Vec3& v) {
    Vec3 r{};
    r[0] = (q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3]) * v[0] +
           2. * (q[1] * q[2] - q[0] * q[3]) * v[1] + 2. * (q[1] * q[3] + q[0] * q[2]) * v[2];
    r[1] = 2. * (q[1] * q[2] + q[0] * q[3]) * v[0] +
           (q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3]) * v[1] +
           2. * (q[2] * q[3] - q[0] * q[1]) * v[2];
    r[2] = 2. * (q[1] * q[3] - q[0] * q[2]) * v[0] + 2. * (q[2] * q[3] + q[0] * q[1]) * v[1] +
           (q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]) * v[2];
    return r;
}
// src/math/quaternion_operations.hpp:152-166 (3x4, row-major)
std::array<double, 12> QuaternionDerivative(const Quat& q) {
    return {-q[1], q[0], -q[3], q[2], -q[2], q[3], q[0], -q[1], -q[3], -q[2], q[1], q[0]};
}
// src/math/quaternion_operations.hpp:171-181
Quat QuaternionInverse(const Quat& q) {
    const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    return {q[0] / len, -q[1] / len, -q[2] / len, -q[3] / len};
}
// src/math/quaternion_operations.hpp:203-211
Quat QuaternionCompose(const Quat& q1, const Quat& q2) {
    Quat qn{};
    qn[0] = q1[0] * q2[0] - q1[1] * q2[1] - q1[2] * q2[2] - q1[3] * q2[3];
    qn[1] = q1[0] * q2[1] + q1[1] * q2[0] + q1[2] * q2[3] - q1[3] * q2[2];
    qn[2] = q1[0] * q2[2] - q1[1] * q2[3] + q1[2] * q2[0] + q1[3] * q2[1];
    qn[3] = q1[0] * q2[3] + q1[1] * q2[2] - q1[2] * q2[1] + q1[3] * q2[0];
    return qn;
}
// src/math/quaternion_operations.hpp:230-242 (device form: factor 0 below 1e-12)
Quat RotationVectorToQuaternion(const Vec3& phi) {
    const double angle = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    const double cos_angle = std::cos(angle / 2.);
    const double factor = (std::abs(angle) < 1.e-12) ? 0. : std::sin(angle / 2.) / angle;
    return {cos_angle, phi[0] * factor, phi[1] * factor, phi[2] * factor};
}
// src/math/quaternion_operations.hpp:287-297 (host std::array form)
Quat RotationVectorToQuaternionHost(const Vec3& phi) {
    const double angle = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    if (std::abs(angle) < 1e-12) return {1., 0., 0., 0.};
    const double s = std::sin(angle / 2.), c = std::cos(angle / 2.);
    const double f = s / angle;
    return {c, phi[0] * f, phi[1] * f, phi[2] * f};
}
// src/math/quaternion_operations.hpp:247-262
Vec3 QuaternionToRotationVector(const Quat& q) {
    const double theta = 2. * std::acos(q[0]);
    const double shh = std::sqrt(1. - q[0] * q[0]);
    if (shh > 1e-12) return {theta * q[1] / shh, theta * q[2] / shh, theta * q[3] / shh};
    return {0., 0., 0.};
}
// src/math/vector_operations.hpp:11-22
Mat3 VecTilde(const Vec3& v) { return {0., -v[2], v[1], v[2], 0., -v[0], -v[1], v[0], 0.}; }
// src/math/matrix_operations.hpp:18-31
Mat3 AX_Matrix(const Mat3& A) {
    double trace = (A[0] + A[4] + A[8]) / 2.;
    Mat3 R{};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = -A[i * 3 + j] / 2.;
        R[i * 3 + i] += trace;
    }
    return R;
}
// src/math/matrix_operations.hpp:44-49
Vec3 AxialVectorOfMatrix(const Mat3& m) {
    return {(m[7] - m[5]) / 2., (m[2] - m[6]) / 2., (m[3] - m[1]) / 2.};
}
// src/math/interpolation.hpp:69-89
void LagrangePolynomialInterpWeights(double x, const std::vector<double>& xs,
                                     std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 1.);
    std::vector<double> xm(n);
    for (size_t m = 0; m < n; ++m) xm[m] = x - xs[m];
    for (size_t j = 0; j < n; ++j)
        for (size_t m = 0; m < n; ++m)
            if (j != m) w[j] *= xm[m] / (xs[j] - xs[m]);
}
// src/math/interpolation.hpp:98-123
void LagrangePolynomialDerivWeights(double x, const std::vector<double>& xs,
                                    std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 0.);
    for (size_t i = 0; i < n; ++i) {
        const double xi = xs[i];
        double weight = 0.;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double prod = 1.;
            for (size_t k = 0; k < n; ++k) {
                if (k == i || k == j) continue;
                prod *= (x - xs[k]) / (xi - xs[k]);
            }
            weight += prod / (xi - xs[j]);
        }
        w[i] = weight;
    }
}
// src/math/interpolation.hpp:21-43
void LinearInterpWeights(double x, const std::vector<double>& xs, std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 0.);
    auto lower = std::lower_bound(xs.begin(), xs.end(), x);
    if (lower == xs.begin()) { w.front() = 1.; return; }
    if (lower == xs.end()) { w.back() = 1.; return; }
    const size_t idx = static_cast<size_t>(lower - xs.begin());
    const double wu = (x - xs[idx - 1]) / (xs[idx] - xs[idx - 1]);
    w[idx - 1] = 1. - wu;
    w[idx] = wu;
}
// src/math/interpolation.hpp:136-157
double LegendrePolynomial(size_t n, double x) {
    if (n == 0) return 1.;
    if (n == 1) return x;
    double p2 = 1., p1 = x, p = 0.;
    for (size_t i = 2; i <= n; ++i) {
        const double d = static_cast<double>(i);
        p = ((2. * d - 1.) * x * p1 - (d - 1.) * p2) / d;
        p2 = p1;
        p1 = p;
    }
    return p;
}
// src/math/interpolation.hpp:169-224
std::vector<double> GenerateGLLPoints(size_t order) {
    if (order < 1) throw std::invalid_argument("Polynomial order must be >= 1");
    const size_t nn = order + 1;
    std::vector<double> pts(nn, 0.);
    pts.front() = -1.;
    pts.back() = 1.;
    std::vector<double> lp(nn, 0.);
    for (size_t i = 1; i < order; ++i) {
        double x = -std::cos(static_cast<double>(i) * kPi / static_cast<double>(order));
        bool converged = false;
        for (int it = 0; it < 1000; ++it) {
            const double xo = x;
            for (size_t k = 0; k < nn; ++k) lp[k] = LegendrePolynomial(k, x);
            const double num = x * lp[nn - 1] - lp[nn - 2];
            const double den = static_cast<double>(nn) * lp[nn - 1];
            x -= num / den;
            if (std::abs(x - xo) <= std::numeric_limits<double>::epsilon()) { converged = true; break; }
        }
        if (!converged) throw std::runtime_error("GLL Newton failed");
        pts[i] = x;
    }
    return pts;
}
// Gauss-Legendre (no reference generator): Newton on P_n with the Tricomi initial guess,
// weights 2 / ((1-x^2) P_n'(x)^2). Pinned by polynomial exactness and the reference's
// 7-point rule (tests/regression_tests/regression/test_cantilever_beam.cpp:74-82).
void GaussLegendreRule(size_t n, std::vector<double>& xi, std::vector<double>& w) {
    xi.assign(n, 0.);
    w.assign(n, 0.);
    for (size_t i = 0; i < n; ++i) {
        const double k = static_cast<double>(i + 1);
        const double nd = static_cast<double>(n);
        double x = std::cos(kPi * (k - 0.25) / (nd + 0.5)) *
                   (1. - (nd - 1.) / (8. * nd * nd * nd));
        double dp = 1.;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1., p1 = x;
            for (size_t m = 2; m <= n; ++m) {
                const double md = static_cast<double>(m);
                const double p2 = ((2. * md - 1.) * x * p1 - (md - 1.) * p0) / md;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1.; }
            dp = nd * (x * p1 - p0) / (x * x - 1.);
            const double dx = p1 / dp;
            x -= dx;
            if (std::abs(dx) < 1e-16) break;
        }
        {
            double p0 = 1., p1 = x;
            for (size_t m = 2; m <= n; ++m) {
                const double md = static_cast<double>(m);
                const double p2 = ((2. * md - 1.) * x * p1 - (md - 1.) * p0) / md;
                p0 = p1;
                p1 = p2;
            }
            dp = nd * (x * p1 - p0) / (x * x - 1.);
        }
        xi[n - 1 - i] = x;  // ascending
        w[n - 1 - i] = 2. / ((1. - x * x) * dp * dp);
    }
    if (n % 2 == 1) xi[n / 2] = 0.;
}

// ---------------------------------------------------------------------------------------
// QP functors
// ---------------------------------------------------------------------------------------
// src/system/masses/rotate_section_matrix.hpp:17-51
Mat6 RotateSectionMatrix(const Quat& xr, const Mat6& Cs) {
    const Mat3 R = QuaternionToRotationMatrix(xr);
    Mat6 T{}, C{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 6; ++j) {
            double a = 0., b = 0.;
            for (int k = 0; k < 3; ++k) {
                a += R[i * 3 + k] * Cs[k * 6 + j];
                b += R[i * 3 + k] * Cs[(3 + k) * 6 + j];
            }
            T[i * 6 + j] = a;
            T[(3 + i) * 6 + j] = b;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0., b = 0.;
            for (int k = 0; k < 3; ++k) {
                a += T[i * 6 + k] * R[j * 3 + k];
                b += T[i * 6 + 3 + k] * R[j * 3 + k];
            }
            C[i * 6 + j] = a;
            C[i * 6 + 3 + j] = b;
        }
    return C;
}
// src/system/beams/calculate_strain.hpp:17-47
std::array<double, 6> CalculateStrain(const Vec3& x0p, const Vec3& up, const Quat& r,
                                      const Quat& rp) {
    Vec3 Rx = RotateVectorByQuaternion(r, x0p);
    for (int i = 0; i < 3; ++i) Rx[i] += -1. * up[i];
    for (int i = 0; i < 3; ++i) Rx[i] += -1. * x0p[i];
    const auto E = QuaternionDerivative(r);
    std::array<double, 3> e2{};
    for (int i = 0; i < 3; ++i) {
        double s = 0.;
        for (int k = 0; k < 4; ++k) s += E[i * 4 + k] * rp[k];
        e2[i] = 2. * s;
    }
    return {-Rx[0], -Rx[1], -Rx[2], e2[0], e2[1], e2[2]};
}
// src/system/beams/calculate_temporary_variables.hpp:17-28
Mat3 CalculateTemporaryVariables(const Vec3& x0p, const Vec3& up) {
    return VecTilde({x0p[0] + up[0], x0p[1] + up[1], x0p[2] + up[2]});
}
// src/system/beams/calculate_force_FC.hpp:15-23
std::array<double, 6> CalculateForceFC(const Mat6& Cuu, const std::array<double, 6>& e) {
    std::array<double, 6> F{};
    for (int i = 0; i < 6; ++i) {
        double s = 0.;
        for (int k = 0; k < 6; ++k) s += Cuu[i * 6 + k] * e[k];
        F[i] = s;
    }
    return F;
}
// src/system/beams/calculate_force_FD.hpp:15-29
std::array<double, 6> CalculateForceFD(const Mat3& x, const std::array<double, 6>& FC) {
    const Vec3 r = mtv(x, {FC[0], FC[1], FC[2]});
    return {0., 0., 0., r[0], r[1], r[2]};
}
// src/system/beams/calculate_Ouu.hpp:17-37
Mat6 CalculateOuu(const Mat6& Cuu, const Mat3& x, const Mat3& Mt, const Mat3& Nt) {
    Mat6 O{};
    const Mat3 a = mm(sub3(Cuu, 0, 0), x), b = mm(sub3(Cuu, 3, 0), x);
    Mat3 o12{}, o22{};
    for (int i = 0; i < 9; ++i) {
        o12[i] = a[i] - Nt[i];
        o22[i] = b[i] - Mt[i];
    }
    put3(O, 0, 3, o12);
    put3(O, 3, 3, o22);
    return O;
}
// src/system/beams/calculate_Puu.hpp:17-36
Mat6 CalculatePuu(const Mat6& Cuu, const Mat3& x, const Mat3& Nt) {
    Mat6 P{};
    const Mat3 a = mtm(x, sub3(Cuu, 0, 0)), b = mtm(x, sub3(Cuu, 0, 3));
    Mat3 p21{};
    for (int i = 0; i < 9; ++i) p21[i] = a[i] + Nt[i];
    put3(P, 3, 0, p21);
    put3(P, 3, 3, b);
    return P;
}
// src/system/beams/calculate_Quu.hpp:17-38
Mat6 CalculateQuu(const Mat6& Cuu, const Mat3& x, const Mat3& Nt) {
    Mat6 Q{};
    const Mat3 a = mm(sub3(Cuu, 0, 0), x);
    Mat3 m1{};
    for (int i = 0; i < 9; ++i) m1[i] = a[i] - Nt[i];
    put3(Q, 3, 3, mtm(x, m1));
    return Q;
}
// src/system/masses/calculate_mass_matrix_components.hpp:8-28
Vec3 CalculateEta(const Mat6& M) {
    const double m = M[0];
    return {M[5 * 6 + 1] / m, -M[5 * 6 + 0] / m, M[4 * 6 + 0] / m};
}
Mat3 CalculateRho(const Mat6& M) { return sub3(M, 3, 3); }
// src/system/masses/calculate_inertial_force.hpp:16-53
std::array<double, 6> CalculateInertialForce(double m, const Vec3& ud, const Vec3& w,
                                             const Vec3& wd, const Vec3& eta, const Mat3& et,
                                             const Mat3& rho, const Mat3& wt, const Mat3& wdt) {
    Mat3 M1 = mm(wt, wt);
    for (int i = 0; i < 9; ++i) M1[i] = m * M1[i] + m * wdt[i];
    const Vec3 f1 = mv(M1, eta);
    std::array<double, 6> F{};
    for (int i = 0; i < 3; ++i) F[i] = f1[i] + m * ud[i];
    const Vec3 v1{m * ud[0], m * ud[1], m * ud[2]};
    const Vec3 a = mv(et, v1), b = mv(rho, wd);
    const Vec3 c = mv(mm(wt, rho), w);
    for (int i = 0; i < 3; ++i) F[3 + i] = a[i] + b[i] + c[i];
    return F;
}
// src/system/masses/calculate_gravity_force.hpp:15-27
std::array<double, 6> CalculateGravityForce(double m, const Vec3& g, const Mat3& et) {
    const Vec3 f{m * g[0], m * g[1], m * g[2]};
    const Vec3 t = mv(et, f);
    return {f[0], f[1], f[2], t[0], t[1], t[2]};
}
// src/system/masses/calculate_gyroscopic_matrix.hpp:20-60
Mat6 CalculateGyroscopicMatrix(double m, const Vec3& w, const Vec3& eta, const Mat3& rho,
                               const Mat3& wt) {
    Mat6 G{};
    const Vec3 v1{m * eta[0], m * eta[1], m * eta[2]};
    const Vec3 v2 = mv(wt, v1);
    const Mat3 a = VecTilde(v2);
    Mat3 g12{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) g12[i * 3 + j] = a[j * 3 + i];
    const Mat3 b = mmt(wt, VecTilde(v1));
    for (int i = 0; i < 9; ++i) g12[i] = b[i] + g12[i];
    put3(G, 0, 3, g12);
    const Mat3 c = VecTilde(mv(rho, w));
    const Mat3 d = mm(wt, rho);
    Mat3 g22{};
    for (int i = 0; i < 9; ++i) g22[i] = d[i] - c[i];
    put3(G, 3, 3, g22);
    return G;
}
// src/system/masses/calculate_inertia_stiffness_matrix.hpp:19-60
Mat6 CalculateInertiaStiffnessMatrix(double m, const Vec3& ud, const Vec3& w, const Vec3& wd,
                                     const Vec3& eta, const Mat3& rho, const Mat3& wt,
                                     const Mat3& wdt) {
    Mat6 K{};
    Mat3 M1 = mm(wt, wt);
    for (int i = 0; i < 9; ++i) M1[i] = M1[i] + wdt[i];
    const Mat3 M2 = VecTilde({m * eta[0], m * eta[1], m * eta[2]});
    put3(K, 0, 3, mmt(M1, M2));
    const Mat3 ut = VecTilde(ud);
    Mat3 k22 = mm(rho, wdt);
    const Mat3 t1 = mm(ut, M2);
    for (int i = 0; i < 9; ++i) k22[i] = t1[i] + k22[i];
    const Mat3 t2 = VecTilde(mv(rho, wd));
    for (int i = 0; i < 9; ++i) k22[i] += -1. * t2[i];
    const Mat3 t3 = VecTilde(mv(rho, w));
    const Mat3 t4 = mm(rho, wt);
    Mat3 m1{};
    for (int i = 0; i < 9; ++i) m1[i] = t4[i] - t3[i];
    const Mat3 t5 = mm(wt, m1);
    for (int i = 0; i < 9; ++i) k22[i] = t5[i] + k22[i];
    put3(K, 3, 3, k22);
    return K;
}
// src/system/calculate_tangent_operator.hpp:26-65 (returns the stored transpose)
Mat6 CalculateTangentOperatorT(double h, const double* qd) {
    Mat6 T{};
    for (int k = 0; k < 6; ++k) T[k * 6 + k] = 1.;
    const Vec3 rv{h * qd[3], h * qd[4], h * qd[5]};
    const double phi = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
    const double tmp1 = (phi > 1.e-16) ? (std::cos(phi) - 1.) / (phi * phi) : 0.;
    const double tmp2 = (phi > 1.e-16) ? (1. - std::sin(phi) / phi) / (phi * phi) : 0.;
    const Mat3 m2 = VecTilde(rv);
    const Mat3 m22 = mm(m2, m2);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            T[(3 + i) * 6 + 3 + j] += tmp2 * m22[i * 3 + j] + tmp1 * m2[i * 3 + j];
    Mat6 Tt{};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Tt[i * 6 + j] = T[j * 6 + i];
    return Tt;
}

// ---------------------------------------------------------------------------------------
// Step parameters — src/step/step_parameters.hpp:37-52
// ---------------------------------------------------------------------------------------
StepParameters::StepParameters(bool dyn, size_t mi, double h_, double rho, double at, double rt)
    : is_dynamic_solve(dyn),
      max_iter(mi),
      h(h_),
      alpha_m((2. * rho - 1.) / (rho + 1.)),
      alpha_f(rho / (rho + 1.)),
      gamma(0.5 + alpha_f - alpha_m),
      beta(0.25 * (gamma + 0.5) * (gamma + 0.5)),
      gamma_prime(dyn ? gamma / (h * beta) : 0.),
      beta_prime(dyn ? (1. - alpha_m) / (h * h * beta * (1. - alpha_f)) : 0.),
      conditioner(beta * h * h),
      absolute_convergence_tol(at),
      relative_convergence_tol(rt) {}

// ---------------------------------------------------------------------------------------
// System setup: model/copy_nodes_to_state.hpp, elements/beams/create_beams.hpp,
// populate_element_views.hpp, calculate_jacobian.hpp, constraints/constraints.hpp,
// dof_management/*.hpp, solver/solver.hpp
// ---------------------------------------------------------------------------------------
// src/elements/beams/calculate_jacobian.hpp:35-76: x0' = sum_k phi'_k x0_k, J = |x0'|, unit
// tangent x0' / J per QP; node_x0 [N][7], shape derivatives [N][Q]
void CalculateJacobian(size_t N, size_t Q, const double* node_x0, const double* dphi, double* J, double* x0p) {
    for (size_t q = 0; q < Q; ++q) {
        std::array<double, 3> d{};
        for (size_t k = 0; k < N; ++k)
            for (int c = 0; c < 3; ++c) d[c] += node_x0[k * 7 + c] * dphi[k * Q + q];
        const double j = std::sqrt(std::pow(d[0], 2.) + std::pow(d[1], 2.) + std::pow(d[2], 2.));
        J[q] = j;
        for (int c = 0; c < 3; ++c) x0p[q * 3 + c] = d[c] / j;
    }
}

// src/state/update_global_position.hpp:24-45
void UpdateGlobalPositionNode(const double* q, const double* x0, double* x) {
    x[0] = x0[0] + q[0];
    x[1] = x0[1] + q[1];
    x[2] = x0[2] + q[2];
    const Quat r = QuaternionCompose({q[3], q[4], q[5], q[6]}, {x0[3], x0[4], x0[5], x0[6]});
    for (int k = 0; k < 4; ++k) x[3 + k] = r[k];
}

namespace {
size_t uf_find(std::vector<size_t>& p, size_t i) {
    while (p[i] != i) {
        p[i] = p[p[i]];
        i = p[i];
    }
    return i;
}
}  // namespace

System CreateSystem(const ModelIn& model) {
    System s;
    const size_t n = model.nodes.size();
    s.num_nodes = n;
    s.x0.assign(n * 7, 0.);
    s.x.assign(n * 7, 0.);
    s.q_prev.assign(n * 7, 0.);
    s.q.assign(n * 7, 0.);
    s.q_delta.assign(n * 6, 0.);
    s.v.assign(n * 6, 0.);
    s.vd.assign(n * 6, 0.);
    s.a.assign(n * 6, 0.);
    s.f.assign(n * 6, 0.);
    s.tangent.assign(n * 36, 0.);
    // src/model/copy_nodes_to_state.hpp:13-60
    for (size_t i = 0; i < n; ++i) {
        const auto& nd = model.nodes[i];
        for (int c = 0; c < 7; ++c) {
            s.x0[i * 7 + c] = nd.x0[c];
            s.q[i * 7 + c] = nd.u[c];
        }
        for (int c = 0; c < 6; ++c) {
            s.v[i * 6 + c] = nd.v[c];
            s.vd[i * 6 + c] = nd.vd[c];
            s.a[i * 6 + c] = nd.vd[c];
        }
    }
    s.q_prev = s.q;
    for (size_t i = 0; i < n; ++i) UpdateGlobalPositionNode(&s.q[i * 7], &s.x0[i * 7], &s.x[i * 7]);
    s.gravity = model.gravity;

    // Beams: src/elements/beams/create_beams.hpp:26-165
    for (const auto& bi : model.beams) {
        BeamElem e;
        e.N = bi.nodes.size();
        e.Q = bi.quad.size();
        const size_t N = e.N, Q = e.Q;
        e.node_idx = bi.nodes;
        e.weight.resize(Q);
        for (size_t q = 0; q < Q; ++q) e.weight[q] = bi.quad[q][1];
        std::vector<double> node_xi(N);
        for (size_t k = 0; k < N; ++k) node_xi[k] = 2. * model.nodes[bi.nodes[k]].s - 1.;
        e.shape_interp.assign(N * Q, 0.);
        e.shape_deriv.assign(N * Q, 0.);
        std::vector<double> w;
        for (size_t q = 0; q < Q; ++q) {
            LagrangePolynomialInterpWeights(bi.quad[q][0], node_xi, w);
            for (size_t k = 0; k < N; ++k) e.shape_interp[k * Q + q] = w[k];
            LagrangePolynomialDerivWeights(bi.quad[q][0], node_xi, w);
            for (size_t k = 0; k < N; ++k) e.shape_deriv[k * Q + q] = w[k];
        }
        // populate_element_views.hpp:89-141 (PopulateQPMstar / PopulateQPCstar)
        std::vector<double> sec_xi(bi.sections.size());
        for (size_t k = 0; k < sec_xi.size(); ++k) sec_xi[k] = 2. * bi.sections[k].s - 1.;
        e.qp_Mstar.assign(Q * 36, 0.);
        e.qp_Cstar.assign(Q * 36, 0.);
        for (size_t q = 0; q < Q; ++q) {
            LinearInterpWeights(bi.quad[q][0], sec_xi, w);
            for (size_t k = 0; k < sec_xi.size(); ++k)
                for (int c = 0; c < 36; ++c) {
                    e.qp_Mstar[q * 36 + c] += bi.sections[k].M[c] * w[k];
                    e.qp_Cstar[q * 36 + c] += bi.sections[k].C[c] * w[k];
                }
        }
        // interpolate_QP_position.hpp, interpolate_QP_rotation.hpp, calculate_jacobian.hpp
        e.qp_x0.assign(Q * 3, 0.);
        e.qp_x0_prime.assign(Q * 3, 0.);
        e.qp_r0.assign(Q * 4, 0.);
        e.jacobian.assign(Q, 0.);
        std::vector<double> node_x0(N * 7);
        for (size_t k = 0; k < N; ++k)
            std::copy(&s.x0[bi.nodes[k] * 7], &s.x0[bi.nodes[k] * 7] + 7, node_x0.begin() + static_cast<long>(k * 7));
        CalculateJacobian(N, Q, node_x0.data(), e.shape_deriv.data(), e.jacobian.data(), e.qp_x0_prime.data());
        for (size_t q = 0; q < Q; ++q) {
            std::array<double, 3> p{};
            std::array<double, 4> r{};
            for (size_t k = 0; k < N; ++k) {
                const double* x0 = &node_x0[k * 7];
                const double phi = e.shape_interp[k * Q + q];
                for (int c = 0; c < 3; ++c) p[c] += x0[c] * phi;
                for (int c = 0; c < 4; ++c) r[c] += x0[3 + c] * phi;
            }
            const double len = std::sqrt(std::pow(r[0], 2) + std::pow(r[1], 2) +
                                         std::pow(r[2], 2) + std::pow(r[3], 2));
            if (len == 0.) r = {1., 0., 0., 0.};
            else
                for (int c = 0; c < 4; ++c) r[c] /= len;
            for (int c = 0; c < 3; ++c) e.qp_x0[q * 3 + c] = p[c];
            for (int c = 0; c < 4; ++c) e.qp_r0[q * 4 + c] = r[c];
        }
        e.qp_Fe.assign(Q * 6, 0.);
        e.node_FX.assign(N * 6, 0.);
        e.residual.assign(N * 6, 0.);
        e.sysmat.assign(N * N * 36, 0.);
        s.beams.push_back(std::move(e));
    }

    // Masses: src/elements/masses/create_masses.hpp:24-60
    for (const auto& mi : model.masses) {
        MassElem e;
        e.node = mi.node;
        e.Mstar = mi.M;
        for (int c = 0; c < 7; ++c) e.x0[c] = model.nodes[mi.node].x0[c];
        s.masses.push_back(e);
    }
    // Springs: src/elements/springs/create_springs.hpp:24-58
    for (const auto& si : model.springs) {
        SpringElem e;
        e.n1 = si.n1;
        e.n2 = si.n2;
        for (int c = 0; c < 3; ++c) e.x0[c] = model.nodes[si.n2].x0[c] - model.nodes[si.n1].x0[c];
        e.l_ref = si.l0;
        e.k = si.k;
        s.springs.push_back(e);
    }

    // Constraints: src/constraints/constraints.hpp:76-266
    size_t start_row = 0;
    for (const auto& ci : model.constraints) {
        Constraint c;
        c.type = ci.type;
        c.base = ci.base;
        c.target = ci.target;
        const bool two = TwoNode(ci.type);
        // freedom signatures and rows (constraints.hpp:181-214, constraint_type.hpp:60-79)
        c.base_sig = two ? kAll : kNone;
        c.target_sig = ThreeDof(ci.type) ? kJustPosition : kAll;
        c.row_count = ThreeDof(ci.type) ? 3 : (ci.type == CType::RevoluteJoint) ? 5 : 6;
        c.row_first = start_row;
        start_row += c.row_count;
        Vec3 X0{0., 0., 0.};
        if (!Prescribed(ci.type)) {
            const auto& t = model.nodes[ci.target].x0;
            if (!two) X0 = {t[0] - ci.axis[0], t[1] - ci.axis[1], t[2] - ci.axis[2]};
            else {
                const auto& b = model.nodes[ci.base].x0;
                X0 = {t[0] - b[0], t[1] - b[1], t[2] - b[2]};
            }
        }
        c.X0 = X0;
        // CalculateAxes (constraints.hpp:296-335)
        Mat3 ax{};
        if (ci.type == CType::RevoluteJoint) {
            const Vec3 xv{1., 0., 0.};
            const double na = std::sqrt(ci.axis[0] * ci.axis[0] + ci.axis[1] * ci.axis[1] +
                                        ci.axis[2] * ci.axis[2]);
            Vec3 xh{};
            if (na != 0.) xh = {ci.axis[0] / na, ci.axis[1] / na, ci.axis[2] / na};
            else {
                const double nx = std::sqrt(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
                if (nx == 0.) throw std::invalid_argument("Cannot normalize a zero vector");
                xh = {X0[0] / nx, X0[1] / nx, X0[2] / nx};
            }
            const Vec3 cp{xv[1] * xh[2] - xv[2] * xh[1], xv[2] * xh[0] - xv[0] * xh[2],
                          xv[0] * xh[1] - xv[1] * xh[0]};
            const double dp = xh[0] * xv[0] + xh[1] * xv[1] + xh[2] * xv[2];
            const double k = 1. / (1. + dp);
            ax = {cp[0] * cp[0] * k + dp, cp[1] * cp[0] * k + cp[2], cp[2] * cp[0] * k - cp[1],
                  cp[0] * cp[1] * k - cp[2], cp[1] * cp[1] * k + dp, cp[2] * cp[1] * k + cp[0],
                  cp[0] * cp[2] * k + cp[1], cp[1] * cp[2] * k - cp[0], cp[2] * cp[2] * k + dp};
        } else if (ci.type == CType::RotationControl) {  // the unit axis (constraints.hpp:324-331)
            const double na = std::sqrt(ci.axis[0] * ci.axis[0] + ci.axis[1] * ci.axis[1] +
                                        ci.axis[2] * ci.axis[2]);
            if (na == 0.) throw std::invalid_argument("Cannot normalize a zero vector");
            ax[0] = ci.axis[0] / na;
            ax[1] = ci.axis[1] / na;
            ax[2] = ci.axis[2] / na;
        } else {
            ax[0] = ci.axis[0];
            ax[1] = ci.axis[1];
            ax[2] = ci.axis[2];
        }
        c.axes = ax;
        if (Prescribed(ci.type)) c.input = ci.init;
        s.cons.push_back(c);
    }

    // DOF tables: assemble_node_freedom_allocation_table.hpp, compute_node_freedom_map_table.hpp
    s.sig.assign(n, kNone);
    for (const auto& e : s.beams)
        for (auto nd : e.node_idx) s.sig[nd] |= kAll;
    for (const auto& e : s.masses) s.sig[e.node] |= kAll;
    for (const auto& e : s.springs) {
        s.sig[e.n1] |= kJustPosition;
        s.sig[e.n2] |= kJustPosition;
    }
    for (const auto& c : s.cons) {
        s.sig[c.target] |= c.target_sig;
        if (TwoNode(c.type)) s.sig[c.base] |= c.base_sig;
    }
    s.active_dofs.assign(n, 0);
    s.dof_map.assign(n, 0);
    size_t acc = 0;
    for (size_t i = 0; i < n; ++i) {
        s.active_dofs[i] = static_cast<size_t>(count_active_dofs(s.sig[i]));
        s.dof_map[i] = acc;
        acc += s.active_dofs[i];
    }
    s.num_system_dofs = acc;
    // create_constraint_freedom_table.hpp
    for (auto& c : s.cons) {
        c.target_dof = s.dof_map[c.target];
        const bool two = TwoNode(c.type);
        c.base_dof = two ? s.dof_map[c.base] : 0;
    }
    s.num_dofs = s.num_system_dofs + start_row;
    s.b.assign(s.num_dofs, 0.);
    s.xsol.assign(s.num_dofs, 0.);

    // Independent blocks (connected components over elements and two-node constraints)
    std::vector<size_t> parent(n);
    std::iota(parent.begin(), parent.end(), size_t{0});
    auto unite = [&](size_t a, size_t b) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    };
    for (const auto& e : s.beams)
        for (size_t k = 1; k < e.N; ++k) unite(e.node_idx[0], e.node_idx[k]);
    for (const auto& e : s.springs) unite(e.n1, e.n2);
    for (const auto& c : s.cons)
        if (TwoNode(c.type)) unite(c.base, c.target);
    std::vector<long> root_to_block(n, -1);
    s.gdof_block.assign(s.num_dofs, 0);
    s.gdof_local.assign(s.num_dofs, 0);
    for (size_t i = 0; i < n; ++i) {
        if (s.active_dofs[i] == 0) continue;
        const size_t r = uf_find(parent, i);
        if (root_to_block[r] < 0) {
            root_to_block[r] = static_cast<long>(s.blocks.size());
            s.blocks.emplace_back();
        }
        auto& blk = s.blocks[static_cast<size_t>(root_to_block[r])];
        for (size_t c = 0; c < s.active_dofs[i]; ++c) blk.gdof.push_back(s.dof_map[i] + c);
    }
    for (const auto& c : s.cons) {
        const size_t r = uf_find(parent, c.target);
        auto& blk = s.blocks[static_cast<size_t>(root_to_block[r])];
        for (size_t k = 0; k < c.row_count; ++k)
            blk.gdof.push_back(s.num_system_dofs + c.row_first + k);
    }
    for (size_t bidx = 0; bidx < s.blocks.size(); ++bidx) {
        auto& blk = s.blocks[bidx];
        blk.n = blk.gdof.size();
        blk.A.assign(blk.n * blk.n, 0.);
        blk.piv.assign(blk.n, 0);
        for (size_t l = 0; l < blk.n; ++l) {
            s.gdof_block[blk.gdof[l]] = bidx;
            s.gdof_local[blk.gdof[l]] = l;
        }
    }
    return s;
}

// ---------------------------------------------------------------------------------------
// Step pieces
// ---------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------
// CRS graph (src/solver/create_full_matrix.hpp:19-74, compute_row_ptrs.hpp:15-71,
// compute_col_inds.hpp:14-61, compute_system_row_entries.hpp:74-102,
// compute_system_col_inds.hpp:90-131, compute_constraints_row_entries.hpp:25-35,
// compute_constraints_col_inds.hpp:49-59, scan_row_entries.hpp)
// ---------------------------------------------------------------------------------------
namespace {
bool ElementContainsNode(const CrsGraphIn& g, size_t e, size_t node) {
    for (size_t k : g.element_nodes[e])
        if (k == node) return true;
    return false;
}
bool ConstraintContainsNode(const CrsGraphIn& g, size_t c, size_t dof_index) {
    return (dof_index == g.base_node_freedom_table[c][0] && g.base_active_dofs[c] != 0) ||
           (dof_index == g.target_node_freedom_table[c][0] && g.target_active_dofs[c] != 0);
}
}  // namespace

std::vector<size_t> ComputeRowPtrs(const CrsGraphIn& g) {
    std::vector<size_t> entries(g.num_dofs, 0);
    for (size_t node = 0; node < g.active_dofs.size(); ++node) {
        const size_t nd = g.active_dofs[node], dof = g.node_freedom_map_table[node];
        size_t row = nd;
        for (size_t e = 0; e < g.element_nodes.size(); ++e) {
            if (!ElementContainsNode(g, e, node)) continue;
            size_t in_elem = 0;
            for (size_t k : g.element_nodes[e]) in_elem += g.active_dofs[k];
            row += in_elem - nd;
        }
        for (size_t c = 0; c < g.row_range.size(); ++c)
            if (ConstraintContainsNode(g, c, dof)) row += g.row_range[c].second - g.row_range[c].first;
        for (size_t k = 0; k < nd; ++k) entries[dof + k] = row;
    }
    for (size_t c = 0; c < g.row_range.size(); ++c)
        for (size_t r = g.row_range[c].first; r < g.row_range[c].second; ++r)
            entries[g.num_system_dofs + r] = g.base_active_dofs[c] + g.target_active_dofs[c];
    std::vector<size_t> row_ptrs(g.num_dofs + 1, 0);
    for (size_t i = 0; i < g.num_dofs; ++i) row_ptrs[i + 1] = row_ptrs[i] + entries[i];
    return row_ptrs;
}

std::vector<size_t> ComputeColInds(const CrsGraphIn& g, const std::vector<size_t>& row_ptrs) {
    std::vector<size_t> col(row_ptrs.back(), 0);
    for (size_t node = 0; node < g.active_dofs.size(); ++node) {
        const size_t nd = g.active_dofs[node], dof = g.node_freedom_map_table[node];
        std::array<size_t, 6> cur{};
        for (size_t c1 = 0; c1 < nd; ++c1) {
            size_t idx = row_ptrs[dof + c1];
            for (size_t c2 = 0; c2 < nd; ++c2) col[idx++] = dof + c2;
            cur[c1] = nd;
        }
        for (size_t e = 0; e < g.element_nodes.size(); ++e) {
            if (!ElementContainsNode(g, e, node)) continue;
            for (size_t c = 0; c < nd; ++c) {
                size_t idx = row_ptrs[dof + c] + cur[c];
                for (size_t k : g.element_nodes[e]) {
                    if (k == node) continue;
                    for (size_t m = 0; m < g.active_dofs[k]; ++m) col[idx++] = g.node_freedom_map_table[k] + m;
                }
                cur[c] = idx - row_ptrs[dof + c];
            }
        }
        for (size_t cn = 0; cn < g.row_range.size(); ++cn) {
            if (!ConstraintContainsNode(g, cn, dof)) continue;
            for (size_t c = 0; c < nd; ++c) {
                size_t idx = row_ptrs[dof + c] + cur[c];
                for (size_t r = g.row_range[cn].first; r < g.row_range[cn].second; ++r) col[idx++] = g.num_system_dofs + r;
                cur[c] = idx - row_ptrs[dof + c];
            }
        }
    }
    for (size_t cn = 0; cn < g.row_range.size(); ++cn)
        for (size_t r = g.row_range[cn].first; r < g.row_range[cn].second; ++r) {
            size_t idx = row_ptrs[r + g.num_system_dofs];
            for (size_t k = 0; k < g.target_active_dofs[cn]; ++k) col[idx++] = g.target_node_freedom_table[cn][k];
            for (size_t k = 0; k < g.base_active_dofs[cn]; ++k) col[idx++] = g.base_node_freedom_table[cn][k];
        }
    return col;
}

// KokkosSparse::sort_crs_graph (create_full_matrix.hpp:62): each row's columns ascending
void SortCrsGraph(const std::vector<size_t>& row_ptrs, std::vector<size_t>& col_inds) {
    for (size_t r = 0; r + 1 < row_ptrs.size(); ++r)
        std::sort(col_inds.begin() + static_cast<long>(row_ptrs[r]), col_inds.begin() + static_cast<long>(row_ptrs[r + 1]));
}

// The graph inputs of a system: Elements' node_state_indices (beams, masses, springs), the
// constraints' base / target freedom tables and row ranges (solver.hpp / constraints.hpp)
CrsGraphIn CrsGraphInput(const System& s) {
    CrsGraphIn g;
    g.num_system_dofs = s.num_system_dofs;
    g.num_dofs = s.num_dofs;
    g.active_dofs = s.active_dofs;
    g.node_freedom_map_table = s.dof_map;
    for (const auto& e : s.beams) g.element_nodes.emplace_back(e.node_idx.begin(), e.node_idx.end());
    for (const auto& m : s.masses) g.element_nodes.push_back({m.node});
    for (const auto& sp : s.springs) g.element_nodes.push_back({sp.n1, sp.n2});
    for (const auto& c : s.cons) {
        const size_t nb = static_cast<size_t>(count_active_dofs(c.base_sig));
        const size_t nt = static_cast<size_t>(count_active_dofs(c.target_sig));
        g.base_active_dofs.push_back(nb);
        g.target_active_dofs.push_back(nt);
        std::array<size_t, 6> bt{}, tt{};
        for (size_t k = 0; k < 6; ++k) {
            bt[k] = c.base_dof + k;
            tt[k] = c.target_dof + k;
        }
        g.base_node_freedom_table.push_back(bt);
        g.target_node_freedom_table.push_back(tt);
        g.row_range.emplace_back(c.row_first, c.row_first + c.row_count);
    }
    return g;
}

// src/state/calculate_displacement.hpp:25-50
void CalculateDisplacementNode(double h, const double* qd, const double* qp, double* q) {
    for (int c = 0; c < 3; ++c) q[c] = qp[c] + h * qd[c];
    const Quat dq = RotationVectorToQuaternion({h * qd[3], h * qd[4], h * qd[5]});
    const Quat r = QuaternionCompose(dq, {qp[3], qp[4], qp[5], qp[6]});
    for (int c = 0; c < 4; ++c) q[3 + c] = r[c];
}
namespace {
double& Aentry(System& s, size_t grow, size_t gcol) {
    const size_t b = s.gdof_block[grow];
    auto& blk = s.blocks[b];
    return blk.A[s.gdof_local[grow] * blk.n + s.gdof_local[gcol]];
}
}  // namespace

// src/step/predict_next_state.hpp:14-47 with src/state/calculate_next_state.hpp:26-37
// src/state/calculate_next_state.hpp:26-37, one node
void CalculateNextStateNode(double h, double alpha_f, double alpha_m, double beta, double gamma, double* qd,
                            double* v, double* vd, double* a) {
    for (int c = 0; c < 6; ++c) {
        const double v_p = v[c], vd_p = vd[c], a_p = a[c];
        vd[c] = 0.;
        a[c] = (alpha_f * vd_p - alpha_m * a_p) / (1. - alpha_m);
        v[c] = v_p + h * (1. - gamma) * a_p + gamma * h * a[c];
        qd[c] = v_p + (0.5 - beta) * h * a_p + beta * h * a[c];
    }
}
void PredictNextState(const StepParameters& p, System& s) {
    s.q_prev = s.q;
    for (size_t i = 0; i < s.num_nodes; ++i)
        CalculateNextStateNode(p.h, p.alpha_f, p.alpha_m, p.beta, p.gamma, &s.q_delta[i * 6], &s.v[i * 6],
                               &s.vd[i * 6], &s.a[i * 6]);
    for (size_t i = 0; i < s.num_nodes; ++i)
        CalculateDisplacementNode(p.h, &s.q_delta[i * 6], &s.q_prev[i * 7], &s.q[i * 7]);
    s.time_step++;
}
// src/step/reset_constraints.hpp:11-15
void ResetConstraints(System& s) {
    for (auto& c : s.cons) {
        c.lambda.fill(0.);
        c.sys_res.fill(0.);
    }
}
// src/step/reset_solver.hpp:11-15
void ResetSolver(System& s) {
    for (auto& blk : s.blocks) std::fill(blk.A.begin(), blk.A.end(), 0.);
    std::fill(s.b.begin(), s.b.end(), 0.);
}
// src/step/update_tangent_operator.hpp:13-26
void UpdateTangentOperator(const StepParameters& p, System& s) {
    for (size_t i = 0; i < s.num_nodes; ++i) {
        const Mat6 T = CalculateTangentOperatorT(p.h, &s.q_delta[i * 6]);
        std::copy(T.begin(), T.end(), s.tangent.begin() + static_cast<long>(i * 36));
    }
}

// src/system/beams/calculate_quadrature_point_values.hpp:50-172 and the functors it calls
void CalculateQuadraturePointValues(double beta_prime, double gamma_prime, const System& s,
                                    BeamElem& e) {
    const size_t N = e.N, Q = e.Q;
    std::vector<double> Fc(Q * 6), Fd(Q * 6), Fi(Q * 6), Fg(Q * 6);
    std::vector<Mat6> Kuu(Q), Puu(Q), Cuu(Q), Ouu(Q), Quu(Q), Muu(Q), Guu(Q);
    for (size_t q = 0; q < Q; ++q) {
        const Quat r0{e.qp_r0[q * 4], e.qp_r0[q * 4 + 1], e.qp_r0[q * 4 + 2], e.qp_r0[q * 4 + 3]};
        Mat6 Ms{}, Cs{};
        std::copy(&e.qp_Mstar[q * 36], &e.qp_Mstar[q * 36] + 36, Ms.begin());
        std::copy(&e.qp_Cstar[q * 36], &e.qp_Cstar[q * 36] + 36, Cs.begin());
        // --- inertial (calculate_inertial_quadrature_point_values.hpp:44-126)
        {
            // interpolate_to_quadrature_point_for_inertia.hpp:19-40
            Quat r{};
            Vec3 ud{}, w{}, wd{};
            for (size_t k = 0; k < N; ++k) {
                const size_t nd = e.node_idx[k];
                const double phi = e.shape_interp[k * Q + q];
                for (int c = 0; c < 3; ++c) {
                    ud[c] += s.vd[nd * 6 + c] * phi;
                    w[c] += s.v[nd * 6 + 3 + c] * phi;
                    wd[c] += s.vd[nd * 6 + 3 + c] * phi;
                }
                for (int c = 0; c < 4; ++c) r[c] += s.q[nd * 7 + 3 + c] * phi;
            }
            const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
            for (int c = 0; c < 4; ++c) r[c] /= rl;
            const Quat xr = QuaternionCompose(r, r0);
            const Mat6 M = RotateSectionMatrix(xr, Ms);
            const double m = M[0];
            const Vec3 eta = CalculateEta(M);
            const Mat3 et = VecTilde(eta);
            const Mat3 rho = CalculateRho(M);
            const Mat3 wt = VecTilde(w), wdt = VecTilde(wd);
            const auto FI = CalculateInertialForce(m, ud, w, wd, eta, et, rho, wt, wdt);
            const auto FG = CalculateGravityForce(m, s.gravity, et);
            Guu[q] = CalculateGyroscopicMatrix(m, w, eta, rho, wt);
            Kuu[q] = CalculateInertiaStiffnessMatrix(m, ud, w, wd, eta, rho, wt, wdt);
            Muu[q] = M;
            for (int c = 0; c < 6; ++c) {
                Fi[q * 6 + c] = FI[c];
                Fg[q * 6 + c] = FG[c];
            }
        }
        // --- stiffness (calculate_stiffness_quadrature_point_values.hpp:47-128)
        {
            // interpolate_to_quadrature_point_for_stiffness.hpp:19-41
            const double J = e.jacobian[q];
            Vec3 u{}, up{};
            Quat r{}, rp{};
            for (size_t k = 0; k < N; ++k) {
                const size_t nd = e.node_idx[k];
                const double phi = e.shape_interp[k * Q + q];
                const double dphiJ = e.shape_deriv[k * Q + q] / J;
                for (int c = 0; c < 3; ++c) {
                    u[c] += s.q[nd * 7 + c] * phi;
                    up[c] += s.q[nd * 7 + c] * dphiJ;
                }
                for (int c = 0; c < 4; ++c) {
                    r[c] += s.q[nd * 7 + 3 + c] * phi;
                    rp[c] += s.q[nd * 7 + 3 + c] * dphiJ;
                }
            }
            const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
            for (int c = 0; c < 4; ++c) r[c] /= rl;
            const Quat xr = QuaternionCompose(r, r0);
            const Mat6 C = RotateSectionMatrix(xr, Cs);
            const Vec3 x0p{e.qp_x0_prime[q * 3], e.qp_x0_prime[q * 3 + 1], e.qp_x0_prime[q * 3 + 2]};
            const auto strain = CalculateStrain(x0p, up, r, rp);
            const Mat3 xs = CalculateTemporaryVariables(x0p, up);
            const auto FC = CalculateForceFC(C, strain);
            const auto FD = CalculateForceFD(xs, FC);
            const Mat3 Nt = VecTilde({FC[0], FC[1], FC[2]});
            const Mat3 Mt = VecTilde({FC[3], FC[4], FC[5]});
            Ouu[q] = CalculateOuu(C, xs, Mt, Nt);
            Puu[q] = CalculatePuu(C, xs, Nt);
            Quu[q] = CalculateQuu(C, xs, Nt);
            Cuu[q] = C;
            for (int c = 0; c < 6; ++c) {
                Fc[q * 6 + c] = FC[c];
                Fd[q * 6 + c] = FD[c];
            }
        }
    }
    IntegrateResidualVector(N, Q, e.weight.data(), e.jacobian.data(), e.shape_interp.data(), e.shape_deriv.data(),
                            Fc.data(), Fd.data(), Fi.data(), e.qp_Fe.data(), Fg.data(), e.node_FX.data(),
                            e.residual.data());
    std::vector<double> K(N * N * 36), M(N * N * 36);
    IntegrateStiffnessMatrix(N, Q, e.weight.data(), e.jacobian.data(), e.shape_interp.data(), e.shape_deriv.data(),
                             Kuu.data(), Puu.data(), Cuu.data(), Ouu.data(), Quu.data(), K.data());
    IntegrateInertiaMatrix(N, Q, e.weight.data(), e.jacobian.data(), e.shape_interp.data(), beta_prime, gamma_prime,
                           Muu.data(), Guu.data(), M.data());
    std::vector<const double*> T(N);
    for (size_t j = 0; j < N; ++j) T[j] = &s.tangent[e.node_idx[j] * 36];
    CalculateSystemMatrix(N, M.data(), K.data(), T.data(), e.sysmat.data());
}

// src/system/beams/integrate_residual_vector.hpp:32-50: R_i = sum_q w (phi'_i Fc + J phi_i (Fd + Fi - Fe - Fg)) - FX_i
void IntegrateResidualVector(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                             const double* dphi, const double* Fc, const double* Fd, const double* Fi,
                             const double* Fe, const double* Fg, const double* node_FX, double* residual) {
    for (size_t k = 0; k < N; ++k) {
        std::array<double, 6> loc{};
        for (size_t q = 0; q < Q; ++q) {
            const double w = weight[q];
            const double cc = w * dphi[k * Q + q];
            const double cd = w * jacobian[q] * phi[k * Q + q];
            for (int c = 0; c < 6; ++c)
                loc[c] += cc * Fc[q * 6 + c] + cd * (Fd[q * 6 + c] + Fi[q * 6 + c] - Fe[q * 6 + c] - Fg[q * 6 + c]);
        }
        for (int c = 0; c < 6; ++c) residual[k * 6 + c] = loc[c] - node_FX[k * 6 + c];
    }
}
// src/system/beams/integrate_stiffness_matrix.hpp:34-96: K_ij = sum_q phi_i phi_j wJ (Kuu + Quu)
// + phi_i phi'_j w Puu + phi'_i phi'_j (w/J) Cuu + phi'_i phi_j w Ouu, out [N][N][36]
void IntegrateStiffnessMatrix(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                              const double* dphi, const Mat6* Kuu, const Mat6* Puu, const Mat6* Cuu, const Mat6* Ouu,
                              const Mat6* Quu, double* out) {
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            Mat6 Ks{};
            for (size_t q = 0; q < Q; ++q) {
                const double w = weight[q], J = jacobian[q];
                const double p1 = phi[i * Q + q], p2 = phi[j * Q + q];
                const double d1 = dphi[i * Q + q], d2 = dphi[j * Q + q];
                const double K = (p1 * p2) * (w * J);
                const double P = (p1 * d2) * w;
                const double C = (d1 * d2) * (w / J);
                const double O = (d1 * p2) * w;
                for (int c = 0; c < 36; ++c)
                    Ks[c] = Ks[c] + K * (Kuu[q][c] + Quu[q][c]) + P * Puu[q][c] + C * Cuu[q][c] + O * Ouu[q][c];
            }
            std::copy(Ks.begin(), Ks.end(), out + (i * N + j) * 36);
        }
}
// src/system/beams/integrate_inertia_matrix.hpp:31-78: M_ij = sum_q phi_i phi_j wJ (beta' Muu + gamma' Guu)
void IntegrateInertiaMatrix(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                            double beta_prime, double gamma_prime, const Mat6* Muu, const Mat6* Guu, double* out) {
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            Mat6 Mi{};
            for (size_t q = 0; q < Q; ++q) {
                const double w = weight[q], J = jacobian[q];
                const double coeff = phi[i * Q + q] * phi[j * Q + q] * w * J;
                for (int c = 0; c < 36; ++c) Mi[c] = Mi[c] + coeff * (beta_prime * Muu[q][c] + gamma_prime * Guu[q][c]);
            }
            std::copy(Mi.begin(), Mi.end(), out + (i * N + j) * 36);
        }
}
// src/system/beams/calculate_system_matrix.hpp:24-53: S_ij = K_ij T_j + M_ij (T_j: node j's stored tangent)
void CalculateSystemMatrix(size_t N, const double* M, const double* K, const double* const* T, double* S) {
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            const double* Kij = K + (i * N + j) * 36;
            const double* Mij = M + (i * N + j) * 36;
            double* Sij = S + (i * N + j) * 36;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    double acc = 0.;
                    for (int k = 0; k < 6; ++k) acc += Kij[r * 6 + k] * T[j][k * 6 + c];
                    Sij[r * 6 + c] = acc + Mij[r * 6 + c];
                }
        }
}

// src/step/update_system_variables_beams.hpp:13-69
void UpdateSystemVariablesBeams(const StepParameters& p, System& s) {
    const long nb = static_cast<long>(s.beams.size());
#pragma omp parallel for schedule(dynamic)
    for (long e = 0; e < nb; ++e)
        CalculateQuadraturePointValues(p.beta_prime, p.gamma_prime, s, s.beams[static_cast<size_t>(e)]);
}
// src/system/masses/calculate_quadrature_point_values.hpp:41-152
void CalculateMassValues(double beta_prime, double gamma_prime, const System& s, MassElem& e) {
    const size_t i = e.node;
    const Vec3 ud{s.vd[i * 6 + 0], s.vd[i * 6 + 1], s.vd[i * 6 + 2]};
    const Vec3 wd{s.vd[i * 6 + 3], s.vd[i * 6 + 4], s.vd[i * 6 + 5]};
    const Vec3 w{s.v[i * 6 + 3], s.v[i * 6 + 4], s.v[i * 6 + 5]};
    const Quat r0{e.x0[3], e.x0[4], e.x0[5], e.x0[6]};
    const Quat r{s.q[i * 7 + 3], s.q[i * 7 + 4], s.q[i * 7 + 5], s.q[i * 7 + 6]};
    const Quat xr = QuaternionCompose(r, r0);
    const Mat3 wt = VecTilde(w), wdt = VecTilde(wd);
    const Mat6 Muu = RotateSectionMatrix(xr, e.Mstar);
    const Vec3 eta = CalculateEta(Muu);
    const Mat3 rho = CalculateRho(Muu);
    const double m = Muu[0];
    const Mat3 et = VecTilde(eta);
    const auto Fg = CalculateGravityForce(m, s.gravity, et);
    const auto Fi = CalculateInertialForce(m, ud, w, wd, eta, et, rho, wt, wdt);
    for (int c = 0; c < 6; ++c) e.residual[c] = Fi[c] - Fg[c];
    const Mat6 Guu = CalculateGyroscopicMatrix(m, w, eta, rho, wt);
    const Mat6 Kuu = CalculateInertiaStiffnessMatrix(m, ud, w, wd, eta, rho, wt, wdt);
    const double* T = &s.tangent[i * 36];
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
            const double v = beta_prime * Muu[a * 6 + b] + gamma_prime * Guu[a * 6 + b];
            // Gemm(1, Kuu, T, 1, STpI): inner product first, then added
            double kt = 0.;
            for (int k = 0; k < 6; ++k) kt += Kuu[a * 6 + k] * T[k * 6 + b];
            e.sysmat[a * 6 + b] = v + kt;
        }
}
// src/system/springs/calculate_quadrature_point_values.hpp:28-90 and the springs/*.hpp helpers
void CalculateSpringValues(const System& s, SpringElem& e) {
    Vec3 r{};
    for (int c = 0; c < 3; ++c) r[c] = e.x0[c] + s.q[e.n2 * 7 + c] - s.q[e.n1 * 7 + c];
    const double l = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double c1 = e.k * (e.l_ref / l - 1.);
    const double c2 = e.k * e.l_ref / (l * l * l);
    const Mat3 rt = VecTilde(r);
    Mat3 a{};
    const double diag = c1 - c2 * l * l;
    const Mat3 rr = mm(rt, rt);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i * 3 + j] = -c2 * rr[i * 3 + j] + (i == j ? diag : 0.);
    for (int c = 0; c < 3; ++c) {
        e.residual[c] = c1 * r[c];
        e.residual[3 + c] = -(c1 * r[c]);
    }
    for (int n1 = 0; n1 < 2; ++n1)
        for (int n2 = 0; n2 < 2; ++n2) {
            const double sg = (n1 == n2) ? 1. : -1.;
            for (int k = 0; k < 9; ++k) e.stiff[(n1 * 2 + n2) * 9 + k] = sg * a[k];
        }
}
void UpdateSystemVariablesMasses(const StepParameters& p, System& s) {
    for (auto& e : s.masses) CalculateMassValues(p.beta_prime, p.gamma_prime, s, e);
}
void UpdateSystemVariablesSprings(System& s) {
    for (auto& e : s.springs) CalculateSpringValues(s, e);
}

// src/step/assemble_system_residual.hpp:17-64 (forces, then beams, masses, springs)
void AssembleSystemResidual(System& s) {
    // contribute_forces_to_vector.hpp:27-48
    for (size_t i = 0; i < s.num_nodes; ++i) {
        const size_t d = s.dof_map[i];
        if (s.sig[i] == kAll)
            for (int c = 0; c < 6; ++c) s.b[d + c] -= s.f[i * 6 + c];
    }
    // contribute_beams_to_vector.hpp:28-47
    for (const auto& e : s.beams)
        for (size_t k = 0; k < e.N; ++k) {
            const size_t d = s.dof_map[e.node_idx[k]];
            for (int c = 0; c < 6; ++c) s.b[d + c] += e.residual[k * 6 + c];
        }
    // contribute_masses_to_vector.hpp
    for (const auto& e : s.masses) {
        const size_t d = s.dof_map[e.node];
        for (int c = 0; c < 6; ++c) s.b[d + c] += e.residual[c];
    }
    // contribute_springs_to_vector.hpp (translational dofs of both nodes)
    for (const auto& e : s.springs) {
        const size_t d1 = s.dof_map[e.n1], d2 = s.dof_map[e.n2];
        for (int c = 0; c < 3; ++c) {
            s.b[d1 + c] += e.residual[c];
            s.b[d2 + c] += e.residual[3 + c];
        }
    }
}
// src/step/assemble_system_matrix.hpp:16-55, contribute_beams_to_sparse_matrix.hpp:34-73
void AssembleSystemMatrix(const StepParameters& p, System& s) {
    for (const auto& e : s.beams)
        for (size_t i = 0; i < e.N; ++i)
            for (size_t j = 0; j < e.N; ++j) {
                const size_t ri = s.dof_map[e.node_idx[i]], cj = s.dof_map[e.node_idx[j]];
                const double* S = &e.sysmat[(i * e.N + j) * 36];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) Aentry(s, ri + r, cj + c) += S[r * 6 + c] * p.conditioner;
            }
    // contribute_masses_to_sparse_matrix.hpp
    for (const auto& e : s.masses) {
        const size_t d = s.dof_map[e.node];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) Aentry(s, d + r, d + c) += e.sysmat[r * 6 + c] * p.conditioner;
    }
    // contribute_springs_to_sparse_matrix.hpp:34-80
    for (const auto& e : s.springs) {
        const size_t d[2] = {s.dof_map[e.n1], s.dof_map[e.n2]};
        for (int n1 = 0; n1 < 2; ++n1)
            for (int n2 = 0; n2 < 2; ++n2)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        Aentry(s, d[n1] + r, d[n2] + c) += e.stiff[(n1 * 2 + n2) * 9 + r * 3 + c] * p.conditioner;
    }
}
// src/step/update_constraint_variables.hpp:13-37 and
// src/constraints/calculate_constraint_residual_gradient.hpp:61-565
void UpdateConstraintVariables(System& s) {
    for (auto& c : s.cons) {
        // Constraints::UpdateViews (constraints.hpp:346-353): the control signal is input(0)
        if (c.type == CType::RevoluteJoint || c.type == CType::RotationControl) c.input[0] = c.control;
        const double* tq = &s.q[c.target * 7];
        const Vec3 u2{tq[0], tq[1], tq[2]};
        const Quat R2{tq[3], tq[4], tq[5], tq[6]};
        Vec3 u1{0., 0., 0.};
        Quat R1{1., 0., 0., 0.};
        if (Prescribed(c.type)) {
            u1 = {c.input[0], c.input[1], c.input[2]};
            R1 = {c.input[3], c.input[4], c.input[5], c.input[6]};
        } else if (TwoNode(c.type)) {
            const double* bq = &s.q[c.base * 7];
            u1 = {bq[0], bq[1], bq[2]};
            R1 = {bq[3], bq[4], bq[5], bq[6]};
        }
        std::array<double, 6> res{};
        Mat6 bg{}, tg{};
        const Quat R1t = QuaternionInverse(R1);
        const Vec3 R1X0 = RotateVectorByQuaternion(R1, c.X0);
        for (int k = 0; k < 3; ++k) res[k] = u2[k] + c.X0[k] - u1[k] - R1X0[k];
        if (ThreeDof(c.type)) {
            // calculate_{fixed_bc_3DOF,prescribed_bc_3DOF,rigid_joint_3DOF}_constraint.hpp: the
            // translational rows only, target gradient I, base gradient [-I tilde(R1 X0)]
            for (int k = 0; k < 3; ++k) tg[k * 6 + k] = 1.;
            if (c.type == CType::RigidJoint6DOFsTo3DOFs) {
                for (int k = 0; k < 3; ++k) bg[k * 6 + k] = -1.;
                put3(bg, 0, 3, VecTilde(R1X0));
            }
        } else if (c.type != CType::RevoluteJoint) {
            // calculate_{fixed_bc,prescribed_bc,rigid_joint}_constraint.hpp; the rotation
            // control (calculate_rotation_control_constraint.hpp:71-136) inserts the inverse of
            // the commanded rotation Rc = exp(axis * input(0)) between target and base
            Quat R2R1t{};
            if (c.type == CType::RotationControl) {
                const Vec3 rv{c.axes[0] * c.input[0], c.axes[1] * c.input[0], c.axes[2] * c.input[0]};
                const Quat RcT = QuaternionInverse(RotationVectorToQuaternion(rv));
                R2R1t = QuaternionCompose(QuaternionCompose(R2, RcT), R1t);
            } else {
                R2R1t = QuaternionCompose(R2, R1t);
            }
            const Mat3 C = QuaternionToRotationMatrix(R2R1t);
            const Vec3 V3 = AxialVectorOfMatrix(C);
            for (int k = 0; k < 3; ++k) res[3 + k] = V3[k];
            for (int k = 0; k < 3; ++k) tg[k * 6 + k] = 1.;
            const Mat3 A = AX_Matrix(C);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) tg[(3 + i) * 6 + 3 + j] = A[j * 3 + i];
            if (c.type == CType::RigidJoint || c.type == CType::RotationControl) {
                for (int k = 0; k < 3; ++k) bg[k * 6 + k] = -1.;
                const Mat3 Xt = VecTilde(R1X0);
                put3(bg, 0, 3, Xt);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) bg[(3 + i) * 6 + 3 + j] = -A[i * 3 + j];
            }
        } else {
            // calculate_revolute_joint_constraint.hpp:20-115
            const Vec3 x0v{c.axes[0], c.axes[1], c.axes[2]};
            const Vec3 y0v{c.axes[3], c.axes[4], c.axes[5]};
            const Vec3 z0v{c.axes[6], c.axes[7], c.axes[8]};
            const Vec3 x = RotateVectorByQuaternion(R1, x0v);
            const Vec3 y = RotateVectorByQuaternion(R2, y0v);
            const Vec3 z = RotateVectorByQuaternion(R2, z0v);
            res[3] = z[0] * x[0] + z[1] * x[1] + z[2] * x[2];
            res[4] = y[0] * x[0] + y[1] * x[1] + y[2] * x[2];
            for (int k = 0; k < 3; ++k) tg[k * 6 + k] = 1.;
            const Vec3 xcz{x[1] * z[2] - x[2] * z[1], x[2] * z[0] - x[0] * z[2], x[0] * z[1] - x[1] * z[0]};
            const Vec3 xcy{x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
            for (int k = 0; k < 3; ++k) {
                tg[3 * 6 + 3 + k] = -xcz[k];
                tg[4 * 6 + 3 + k] = -xcy[k];
            }
            for (int k = 0; k < 3; ++k) bg[k * 6 + k] = -1.;
            for (int k = 0; k < 3; ++k) {
                bg[3 * 6 + 3 + k] = xcz[k];
                bg[4 * 6 + 3 + k] = xcy[k];
            }
            // calculate_revolute_joint_force.hpp:20-47
            const Vec3 R2X0 = RotateVectorByQuaternion(R2, x0v);
            c.sys_res[3] = R2X0[0] * c.input[0];
            c.sys_res[4] = R2X0[1] * c.input[0];
            c.sys_res[5] = R2X0[2] * c.input[0];
        }
        // transpose, B*tangent, B^T*lambda (calculate_constraint_residual_gradient.hpp:88-99)
        auto finish = [&](const Mat6& g, const double* tan, Mat6& gtan, Mat6& gT,
                          std::array<double, 6>& lres) {
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) gT[i * 6 + j] = g[j * 6 + i];
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    double acc = 0.;
                    for (int k = 0; k < 6; ++k) acc += g[i * 6 + k] * tan[k * 6 + j];
                    gtan[i * 6 + j] = acc;
                }
            for (int i = 0; i < 6; ++i) {
                double acc = 0.;
                for (int k = 0; k < 6; ++k) acc += gT[i * 6 + k] * c.lambda[k];
                lres[i] = acc;
            }
        };
        finish(tg, &s.tangent[c.target * 36], c.t_grad, c.t_grad_T, c.t_lres);
        if (TwoNode(c.type))
            finish(bg, &s.tangent[c.base * 36], c.b_grad, c.b_grad_T, c.b_lres);
        c.res = res;
    }
}
// src/step/assemble_constraints_matrix.hpp:13-42, copy_constraints_to_sparse_matrix.hpp:37-84,
// copy_constraints_transpose_to_sparse_matrix.hpp:37-98
void AssembleConstraintsMatrix(System& s) {
    for (const auto& c : s.cons) {
        const int nb = count_active_dofs(c.base_sig), nt = count_active_dofs(c.target_sig);
        for (size_t i = 0; i < c.row_count; ++i) {
            const size_t row = s.num_system_dofs + c.row_first + i;
            for (int k = 0; k < nb; ++k) Aentry(s, row, c.base_dof + k) += c.b_grad[i * 6 + k];
            for (int k = 0; k < nt; ++k) Aentry(s, row, c.target_dof + k) += c.t_grad[i * 6 + k];
        }
        const size_t col0 = s.num_system_dofs + c.row_first;
        for (int r = 0; r < nb; ++r)
            for (size_t k = 0; k < c.row_count; ++k) Aentry(s, c.base_dof + r, col0 + k) += c.b_grad_T[r * 6 + k];
        for (int r = 0; r < nt; ++r)
            for (size_t k = 0; k < c.row_count; ++k) Aentry(s, c.target_dof + r, col0 + k) += c.t_grad_T[r * 6 + k];
    }
}
// src/step/assemble_constraints_residual.hpp:15-51
void AssembleConstraintsResidual(System& s) {
    if (s.cons.empty()) return;
    for (const auto& c : s.cons) {
        const int nt = count_active_dofs(c.target_sig);
        for (int k = 0; k < nt; ++k) s.b[c.target_dof + k] += c.sys_res[k];
    }
    for (const auto& c : s.cons) {
        const int nb = count_active_dofs(c.base_sig), nt = count_active_dofs(c.target_sig);
        for (int k = 0; k < nb; ++k) s.b[c.base_dof + k] += c.b_lres[k];
        for (int k = 0; k < nt; ++k) s.b[c.target_dof + k] += c.t_lres[k];
    }
    for (const auto& c : s.cons)
        for (size_t k = 0; k < c.row_count; ++k) s.b[s.num_system_dofs + c.row_first + k] += c.res[k];
}

bool LUFactor(size_t n, double* A, int* piv) {
    bool ok = true;
    for (size_t k = 0; k < n; ++k) {
        size_t p = k;
        double pmax = std::abs(A[k * n + k]);
        for (size_t i = k + 1; i < n; ++i) {
            const double v = std::abs(A[i * n + k]);
            if (v > pmax) {
                pmax = v;
                p = i;
            }
        }
        piv[k] = static_cast<int>(p);
        if (p != k)
            for (size_t j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
        const double d = A[k * n + k];
        if (d == 0.) {
            ok = false;
            continue;
        }
        const double inv = 1. / d;
        for (size_t i = k + 1; i < n; ++i) {
            const double l = A[i * n + k] * inv;
            A[i * n + k] = l;
            for (size_t j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
        }
    }
    return ok;
}
void LUSolve(size_t n, const double* A, const int* piv, double* b) {
    for (size_t k = 0; k < n; ++k) {
        const size_t p = static_cast<size_t>(piv[k]);
        if (p != k) std::swap(b[k], b[p]);
    }
    for (size_t i = 0; i < n; ++i) {
        double acc = b[i];
        for (size_t j = 0; j < i; ++j) acc -= A[i * n + j] * b[j];
        b[i] = acc;
    }
    // column-oriented back substitution (x_k final, then eliminated from the rows above)
    for (size_t k = n; k-- > 0;) {
        b[k] = b[k] / A[k * n + k];
        for (size_t i = 0; i < k; ++i) b[i] -= A[i * n + k] * b[k];
    }
}

// src/step/solve_system.hpp:14-39 with src/solver/condition_system.hpp:10-31
void SolveSystem(const StepParameters& p, System& s) {
    for (size_t i = 0; i < s.num_system_dofs; ++i) s.b[i] *= p.conditioner;
    for (auto& v : s.b) v *= -1.;
    const long nb = static_cast<long>(s.blocks.size());
#pragma omp parallel for schedule(dynamic)
    for (long bi = 0; bi < nb; ++bi) {
        auto& blk = s.blocks[static_cast<size_t>(bi)];
        std::vector<double> rhs(blk.n);
        for (size_t l = 0; l < blk.n; ++l) rhs[l] = s.b[blk.gdof[l]];
        LUFactor(blk.n, blk.A.data(), blk.piv.data());
        LUSolve(blk.n, blk.A.data(), blk.piv.data(), rhs.data());
        for (size_t l = 0; l < blk.n; ++l) s.xsol[blk.gdof[l]] = rhs[l];
    }
    for (size_t i = s.num_system_dofs; i < s.num_dofs; ++i) s.xsol[i] /= p.conditioner;
}
// src/step/calculate_convergence_error.hpp:16-54, src/solver/calculate_error_sum_squares.hpp
double CalculateConvergenceError(const StepParameters& p, const System& s) {
    const double at = p.absolute_convergence_tol, rt = p.relative_convergence_tol;
    double sys = 0.;
    for (size_t i = 0; i < s.num_nodes; ++i)
        for (size_t c = 0; c < s.active_dofs[i]; ++c) {
            const double pi = s.xsol[s.dof_map[i] + c];
            const double xi = s.q_delta[i * 6 + c] * p.h;
            const double e = pi / (at + rt * std::abs(xi));
            sys += e * e;
        }
    double con = 0.;
    for (const auto& c : s.cons)
        for (size_t r = 0; r < c.row_count; ++r) {
            const double pi = s.xsol[s.num_system_dofs + c.row_first + r];
            const double xi = c.lambda[r];
            const double e = pi / (at + rt * std::abs(xi));
            con += e * e;
        }
    return std::sqrt((sys + con) / static_cast<double>(s.num_dofs));
}
// src/step/update_state_prediction.hpp:25-68, src/state/update_{dynamic,static}_prediction.hpp
// src/state/update_dynamic_prediction.hpp:34-44, one node with `nd` active dofs
void UpdateDynamicPredictionNode(double h, double beta_prime, double gamma_prime, const double* x, size_t nd,
                                 double* qd, double* v, double* vd) {
    for (size_t c = 0; c < nd; ++c) {
        const double dx = x[c];
        qd[c] += dx / h;
        v[c] += gamma_prime * dx;
        vd[c] += beta_prime * dx;
    }
}
// src/state/update_static_prediction.hpp:12-38, one node with `nd` active dofs
void UpdateStaticPredictionNode(double h, const double* x, size_t nd, double* qd) {
    for (size_t c = 0; c < nd; ++c) qd[c] += x[c] / h;
}
void UpdateStatePrediction(const StepParameters& p, System& s) {
    for (size_t i = 0; i < s.num_nodes; ++i) {
        const size_t nd = s.active_dofs[i], d0 = s.dof_map[i];
        if (p.is_dynamic_solve)
            UpdateDynamicPredictionNode(p.h, p.beta_prime, p.gamma_prime, &s.xsol[d0], nd, &s.q_delta[i * 6],
                                        &s.v[i * 6], &s.vd[i * 6]);
        else
            UpdateStaticPredictionNode(p.h, &s.xsol[d0], nd, &s.q_delta[i * 6]);
    }
    for (size_t i = 0; i < s.num_nodes; ++i)
        CalculateDisplacementNode(p.h, &s.q_delta[i * 6], &s.q_prev[i * 7], &s.q[i * 7]);
}
// src/step/update_constraint_prediction.hpp:13-25, src/constraints/update_lambda_prediction.hpp
void UpdateConstraintPrediction(System& s) {
    for (auto& c : s.cons)
        for (size_t r = 0; r < c.row_count; ++r) c.lambda[r] += s.xsol[s.num_system_dofs + c.row_first + r];
}
// src/step/step.hpp:85-121
// src/state/update_algorithmic_acceleration.hpp:18-23, one node
void UpdateAlgorithmicAccelerationNode(double alpha_f, double alpha_m, const double* vd, double* a) {
    for (int c = 0; c < 6; ++c) a[c] += (1. - alpha_f) / (1. - alpha_m) * vd[c];
}
void StepEpilogue(const StepParameters& p, System& s) {
    for (size_t i = 0; i < s.num_nodes; ++i) UpdateAlgorithmicAccelerationNode(p.alpha_f, p.alpha_m, &s.vd[i * 6], &s.a[i * 6]);
    for (size_t i = 0; i < s.num_nodes; ++i) UpdateGlobalPositionNode(&s.q[i * 7], &s.x0[i * 7], &s.x[i * 7]);
    // calculate_constraint_output.hpp, calculate_revolute_joint_output.hpp
    for (auto& c : s.cons) {
        if (c.type != CType::RevoluteJoint) continue;
        const double* q = &s.q[c.target * 7];
        const Quat R{q[3], q[4], q[5], q[6]};
        const Vec3 rv = QuaternionToRotationVector(R);
        const Vec3 ax = RotateVectorByQuaternion(R, {c.axes[0], c.axes[1], c.axes[2]});
        const double* v = &s.v[c.target * 6];
        const double* vd = &s.vd[c.target * 6];
        c.output[0] = ax[0] * rv[0] + ax[1] * rv[1] + ax[2] * rv[2];
        c.output[1] = ax[0] * v[3] + ax[1] * v[4] + ax[2] * v[5];
        c.output[2] = ax[0] * vd[3] + ax[1] * vd[4] + ax[2] * vd[5];
    }
}
// src/step/step.hpp:42-124
bool Step(const StepParameters& p, System& s) {
    PredictNextState(p, s);
    ResetConstraints(s);
    s.convergence_err.clear();
    double err = 1000.;
    for (size_t iter = 0; err > 1.; ++iter) {
        if (iter >= p.max_iter) return false;
        ResetSolver(s);
        UpdateTangentOperator(p, s);
        UpdateSystemVariablesBeams(p, s);
        UpdateSystemVariablesMasses(p, s);
        UpdateSystemVariablesSprings(s);
        AssembleSystemResidual(s);
        AssembleSystemMatrix(p, s);
        UpdateConstraintVariables(s);
        AssembleConstraintsMatrix(s);
        AssembleConstraintsResidual(s);
        SolveSystem(p, s);
        err = CalculateConvergenceError(p, s);
        s.convergence_err.push_back(err);
        UpdateStatePrediction(p, s);
        UpdateConstraintPrediction(s);
    }
    StepEpilogue(p, s);
    return true;
}

}  // namespace kyn_oracle
