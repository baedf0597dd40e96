// Blade interface over the C-ABI: the Beam component builder (host setup) and the
// BladeInterface step loop (loads in, one device step, packed node motion out).
//
//   components::Beam::Beam            src/interfaces/components/beam.cpp:17-24
//     SetupNodeLocations / CreateNodeGeometry / CalcNodeTangents   beam.cpp:76-184
//     CreateBeamElement / BuildBeamSections                        beam.cpp:112-281
//     PositionBladeInSpace / SetupRootNode                         beam.cpp:139-172
//   BeamBuilder::AddRefAxisPoint / AddSection (Z-axis inputs)      beam_builder.cpp:26-90
//   interfaces::BladeInterface        src/interfaces/blade/blade_interface.cpp:11-103
//
// Math restated from src/math/interpolation.hpp (GLL points, Lagrange weights, linear
// interpolation), least_squares_fit.hpp, project_points_to_target_polynomial.hpp,
// quaternion_operations.hpp (RotationMatrixToQuaternion, TangentTwistToQuaternion),
// matrix_operations.hpp (RotateMatrix6), elements/beams/beam_quadrature.hpp and
// model/node.hpp (Translate, RotateAboutPoint, Set{Velocity,Acceleration}AboutPoint).
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kynema_amd.h"

// the library's thread-local error message (kyn_last_error), kept in kyn_host.cpp
namespace kyn {
void set_last_error(const std::string& msg);
}

namespace {

using V3 = std::array<double, 3>;
using Q4 = std::array<double, 4>;
using M6 = std::array<double, 36>;
constexpr double kPi = 3.14159265358979323846;
constexpr size_t kMaxGeometryPoints = 10;  // beam.hpp:33

struct BladeError : std::runtime_error {
    int code;
    BladeError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---- quaternion / vector helpers (quaternion_operations.hpp, vector_operations.hpp)
Q4 compose(const Q4& a, const Q4& b) {
    return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3], a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
            a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1], a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}
std::array<std::array<double, 3>, 3> to_rot(const Q4& q) {
    return {{{q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3], 2. * (q[1] * q[2] - q[0] * q[3]),
              2. * (q[1] * q[3] + q[0] * q[2])},
             {2. * (q[1] * q[2] + q[0] * q[3]), q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3],
              2. * (q[2] * q[3] - q[0] * q[1])},
             {2. * (q[1] * q[3] - q[0] * q[2]), 2. * (q[2] * q[3] + q[0] * q[1]),
              q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]}}};
}
V3 rotate(const Q4& q, const V3& v) {
    const auto R = to_rot(q);
    return {R[0][0] * v[0] + R[0][1] * v[1] + R[0][2] * v[2], R[1][0] * v[0] + R[1][1] * v[1] + R[1][2] * v[2],
            R[2][0] * v[0] + R[2][1] * v[1] + R[2][2] * v[2]};
}
Q4 rotvec_to_quat(const V3& phi) {  // RotationVectorToQuaternion (quaternion_operations.hpp:230-242)
    const double angle = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    if (std::abs(angle) < 1e-12) return {1., 0., 0., 0.};
    const double sa = std::sin(angle / 2.), ca = std::cos(angle / 2.);
    return {ca, phi[0] / angle * sa, phi[1] / angle * sa, phi[2] / angle * sa};
}
V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 unit(const V3& a) {
    const double n = std::sqrt(dot(a, a));
    if (n == 0.) throw BladeError(KYN_ERR_INVALID, "Cannot normalize a zero vector");
    return {a[0] / n, a[1] / n, a[2] / n};
}
// RotationMatrixToQuaternion (quaternion_operations.hpp:63-113): the largest of the four
// trace combinations selects the stable formula
Q4 rot_to_quat(const std::array<std::array<double, 3>, 3>& m) {
    const double p = m[1][1] + m[2][2], d = m[1][1] - m[2][2];
    const std::array<double, 4> vals{m[0][0] + p, m[0][0] - p, -m[0][0] + d, -m[0][0] - d};
    const auto idx = std::distance(vals.begin(), std::max_element(vals.begin(), vals.end()));
    const double tmp = std::sqrt(vals[idx] + 1.), c = 0.5 / tmp;
    switch (idx) {
        case 0: return {0.5 * tmp, (m[2][1] - m[1][2]) * c, (m[0][2] - m[2][0]) * c, (m[1][0] - m[0][1]) * c};
        case 1: return {(m[2][1] - m[1][2]) * c, 0.5 * tmp, (m[0][1] + m[1][0]) * c, (m[0][2] + m[2][0]) * c};
        case 2: return {(m[0][2] - m[2][0]) * c, (m[0][1] + m[1][0]) * c, 0.5 * tmp, (m[1][2] + m[2][1]) * c};
        default: return {(m[1][0] - m[0][1]) * c, (m[0][2] + m[2][0]) * c, (m[1][2] + m[2][1]) * c, 0.5 * tmp};
    }
}
// TangentTwistToQuaternion (quaternion_operations.hpp:334-362)
Q4 tangent_twist_to_quat(const V3& tangent, double twist_deg) {
    const V3 e1 = unit(tangent);
    V3 t{0., 1., 0.};
    if (std::abs(dot(e1, t)) > 0.9) t = {1., 0., 0.};
    const double a = dot(e1, t);
    const V3 e2 = unit({t[0] - e1[0] * a, t[1] - e1[1] * a, t[2] - e1[2] * a});
    const V3 e3 = cross(e1, e2);
    const Q4 q_tan = rot_to_quat({{{e1[0], e2[0], e3[0]}, {e1[1], e2[1], e3[1]}, {e1[2], e2[2], e3[2]}}});
    const double tr = twist_deg * kPi / 180.;
    return compose(rotvec_to_quat({e1[0] * tr, e1[1] * tr, e1[2] * tr}), q_tan);
}
// RotateMatrix6 (matrix_operations.hpp:53-90): blkdiag(R, R) m blkdiag(R, R)^T
M6 rotate_matrix6(const M6& m, const Q4& q) {
    const auto R = to_rot(q);
    double r[6][6] = {};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = r[3 + i][3 + j] = R[i][j];
    double mt[6][6];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            mt[i][j] = 0.;
            for (int k = 0; k < 6; ++k) mt[i][j] += r[i][k] * m[k * 6 + j];
        }
    M6 out{};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double acc = 0.;
            for (int k = 0; k < 6; ++k) acc += mt[i][k] * r[j][k];
            out[i * 6 + j] = acc;
        }
    return out;
}

// ---- interpolation (interpolation.hpp)
double legendre(size_t n, double x) {
    if (n == 0) return 1.;
    if (n == 1) return x;
    double p2 = 1., p1 = x, p = 0.;
    for (size_t i = 2; i <= n; ++i) {
        const double di = static_cast<double>(i);
        p = ((2. * di - 1.) * x * p1 - (di - 1.) * p2) / di;
        p2 = p1;
        p1 = p;
    }
    return p;
}
std::vector<double> gll_points(size_t order) {  // GenerateGLLPoints (:169-224)
    if (order < 1) throw BladeError(KYN_ERR_INVALID, "Polynomial order must be >= 1");
    const size_t n = order + 1;
    std::vector<double> pts(n, 0.), P(n, 0.);
    pts.front() = -1.;
    pts.back() = 1.;
    for (size_t i = 1; i < order; ++i) {
        double x = -std::cos(static_cast<double>(i) * kPi / static_cast<double>(order));
        bool ok = false;
        for (int it = 0; it < 1000; ++it) {
            const double xo = x;
            for (size_t k = 0; k < n; ++k) P[k] = legendre(k, x);
            x -= (x * P[n - 1] - P[n - 2]) / (static_cast<double>(n) * P[n - 1]);
            if (std::abs(x - xo) <= std::numeric_limits<double>::epsilon()) {
                ok = true;
                break;
            }
        }
        if (!ok) throw BladeError(KYN_ERR_INVALID, "GLL Newton iteration did not converge");
        pts[i] = x;
    }
    return pts;
}
void lagrange_weights(double x, const std::vector<double>& xs, std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 1.);
    for (size_t j = 0; j < n; ++j)
        for (size_t m = 0; m < n; ++m)
            if (m != j) w[j] *= (x - xs[m]) / (xs[j] - xs[m]);
}
void lagrange_deriv_weights(double x, const std::vector<double>& xs, std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 0.);
    for (size_t i = 0; i < n; ++i) {
        double acc = 0.;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double prod = 1.;
            for (size_t k = 0; k < n; ++k)
                if (k != i && k != j) prod *= (x - xs[k]) / (xs[i] - xs[k]);
            acc += prod / (xs[i] - xs[j]);
        }
        w[i] = acc;
    }
}
// [outputs][inputs]: basis on `outputs` evaluated at each input point (least_squares_fit.hpp)
std::vector<std::vector<double>> shape_values(const std::vector<double>& in, const std::vector<double>& out) {
    std::vector<std::vector<double>> S(out.size(), std::vector<double>(in.size(), 0.));
    std::vector<double> w;
    for (size_t i = 0; i < in.size(); ++i) {
        lagrange_weights(in[i], out, w);
        for (size_t o = 0; o < out.size(); ++o) S[o][i] = w[o];
    }
    return S;
}
std::vector<std::vector<double>> shape_derivs(const std::vector<double>& in, const std::vector<double>& out) {
    std::vector<std::vector<double>> S(out.size(), std::vector<double>(in.size(), 0.));
    std::vector<double> w;
    for (size_t i = 0; i < in.size(); ++i) {
        lagrange_deriv_weights(in[i], out, w);
        for (size_t o = 0; o < out.size(); ++o) S[o][i] = w[o];
    }
    return S;
}
double linear_interp(double x, const std::vector<double>& xs, const std::vector<double>& vals) {
    const auto lower = std::lower_bound(xs.begin(), xs.end(), x);
    if (lower == xs.begin()) return vals.front();
    if (lower == xs.end()) return vals.back();
    const auto i = static_cast<size_t>(std::distance(xs.begin(), lower));
    const double wu = (x - xs[i - 1]) / (xs[i] - xs[i - 1]);
    return (1. - wu) * vals[i - 1] + wu * vals[i];
}

// PerformLeastSquaresFitting (least_squares_fit.hpp:126-190): end points pinned, normal
// equations for the interior; the dense p x p solve by Gaussian elimination with partial
// pivoting (dgesv's algorithm)
std::vector<V3> least_squares(size_t p, const std::vector<std::vector<double>>& phi, const std::vector<V3>& pts) {
    const size_t n = phi[0].size();
    std::vector<double> A(p * p, 0.), B(p * 3, 0.);
    A[0] = 1.;
    A[(p - 1) * p + p - 1] = 1.;
    for (size_t j = 0; j < p; ++j)
        for (size_t i = 1; i + 1 < p; ++i) {
            double acc = 0.;
            for (size_t k = 0; k < n; ++k) acc += phi[i][k] * phi[j][k];
            A[i * p + j] = acc;
        }
    for (int d = 0; d < 3; ++d) {
        B[d] = pts[0][d];
        B[(p - 1) * 3 + d] = pts[n - 1][d];
    }
    for (size_t i = 1; i + 1 < p; ++i)
        for (size_t k = 0; k < n; ++k)
            for (int d = 0; d < 3; ++d) B[i * 3 + d] += phi[i][k] * pts[k][d];
    for (size_t k = 0; k < p; ++k) {
        size_t piv = k;
        for (size_t i = k + 1; i < p; ++i)
            if (std::abs(A[i * p + k]) > std::abs(A[piv * p + k])) piv = i;
        if (piv != k) {
            for (size_t j = 0; j < p; ++j) std::swap(A[k * p + j], A[piv * p + j]);
            for (int d = 0; d < 3; ++d) std::swap(B[k * 3 + d], B[piv * 3 + d]);
        }
        if (A[k * p + k] == 0.) throw BladeError(KYN_ERR_INVALID, "singular least-squares system");
        for (size_t i = k + 1; i < p; ++i) {
            const double l = A[i * p + k] / A[k * p + k];
            for (size_t j = k; j < p; ++j) A[i * p + j] -= l * A[k * p + j];
            for (int d = 0; d < 3; ++d) B[i * 3 + d] -= l * B[k * 3 + d];
        }
    }
    std::vector<V3> X(p);
    for (size_t k = p; k-- > 0;)
        for (int d = 0; d < 3; ++d) {
            double acc = B[k * 3 + d];
            for (size_t j = k + 1; j < p; ++j) acc -= A[k * p + j] * X[j][d];
            X[k][d] = acc / A[k * p + k];
        }
    return X;
}

// CreateTrapezoidalQuadrature (beam_quadrature.hpp:11-28)
std::vector<double> trapezoidal_quadrature(const std::vector<double>& g) {
    const size_t n = g.size();
    const double lo = *std::min_element(g.begin(), g.end()), hi = *std::max_element(g.begin(), g.end());
    const double range = hi - lo;
    std::vector<double> q{-1., (g[1] - g[0]) / range};
    for (size_t i = 1; i + 1 < n; ++i) {
        q.push_back(2. * (g[i] - lo) / range - 1.);
        q.push_back((g[i + 1] - g[i - 1]) / range);
    }
    q.push_back(1.);
    q.push_back((g[n - 1] - g[n - 2]) / range);
    return q;
}

struct Built {
    std::vector<double> node_xi;
    std::vector<size_t> node_ids;
    int64_t beam = -1, root_constraint = -1;
};

// components::Beam::Beam (beam.cpp:17-24)
Built build_beam(kyn_model m, const kyn_beam_input& in) {
    if (in.num_ref_axis_points < 2 || !in.ref_axis_grid || !in.ref_axis_coordinates)
        throw BladeError(KYN_ERR_INVALID, "At least two reference axis points are required");
    if (in.num_sections == 0 || !in.section_grid || !in.section_mass || !in.section_stiffness)
        throw BladeError(KYN_ERR_INVALID, "At least one section is required");
    if (in.element_order < 1)
        throw BladeError(KYN_ERR_INVALID, "Element order must be at least 1 i.e. linear element for discretization");
    if (in.num_twist_points < 1 || !in.twist_grid || !in.twist)
        throw BladeError(KYN_ERR_INVALID, "At least one reference axis twist point is required");
    const bool z_axis = in.reference_axis_orientation == KYN_REF_AXIS_Z;
    if (!z_axis && in.reference_axis_orientation != KYN_REF_AXIS_X)
        throw BladeError(KYN_ERR_INVALID, "Invalid reference axis orientation");
    const Q4 q_z_to_x = rotvec_to_quat({0., kPi / 2., 0.});  // beam_builder.cpp:35-38

    std::vector<double> grid(in.ref_axis_grid, in.ref_axis_grid + in.num_ref_axis_points);
    std::vector<V3> coords(in.num_ref_axis_points);
    for (size_t k = 0; k < coords.size(); ++k) {
        const V3 c{in.ref_axis_coordinates[3 * k], in.ref_axis_coordinates[3 * k + 1], in.ref_axis_coordinates[3 * k + 2]};
        coords[k] = z_axis ? rotate(q_z_to_x, c) : c;
    }
    Built B;
    B.node_xi = gll_points(in.element_order);  // SetupNodeLocations
    const size_t n_nodes = in.element_order + 1;
    // CreateNodeGeometry (beam.cpp:81-108)
    const size_t n_geo = std::min({grid.size(), n_nodes, kMaxGeometryPoints});
    if (grid.back() == grid.front())
        throw BladeError(KYN_ERR_INVALID, "Invalid geometric locations: domain start and end points are equal.");
    std::vector<double> kp_xi(grid.size());
    for (size_t k = 0; k < grid.size(); ++k) kp_xi[k] = 2. * (grid[k] - grid.front()) / (grid.back() - grid.front()) - 1.;
    std::vector<V3> node_coords(n_nodes, V3{0., 0., 0.});
    if (n_geo < n_nodes) {
        const auto gll_geo = gll_points(n_geo - 1);
        const auto geo = least_squares(n_geo, shape_values(kp_xi, gll_geo), coords);
        // ProjectPointsToTargetPolynomial (project_points_to_target_polynomial.hpp:17-49)
        const auto proj = shape_values(gll_points(n_nodes - 1), gll_points(n_geo - 1));  // [n_geo][n_nodes]
        for (size_t o = 0; o < n_nodes; ++o)
            for (size_t i = 0; i < n_geo; ++i)
                for (int d = 0; d < 3; ++d) node_coords[o][d] += proj[i][o] * geo[i][d];
    } else {
        node_coords = least_squares(n_nodes, shape_values(kp_xi, B.node_xi), coords);
    }
    // CalcNodeTangents (beam.cpp:174-208)
    const auto dphi = shape_derivs(B.node_xi, B.node_xi);
    std::vector<V3> tangents(n_nodes, V3{0., 0., 0.});
    for (size_t n1 = 0; n1 < n_nodes; ++n1)
        for (size_t n2 = 0; n2 < n_nodes; ++n2)
            for (int d = 0; d < 3; ++d) tangents[n2][d] += dphi[n1][n2] * node_coords[n1][d];
    for (auto& t : tangents) t = unit(t);

    // nodes, placed in space (PositionBladeInSpace, beam.cpp:139-164; node.hpp)
    const V3 root_loc{in.root_position[0], in.root_position[1], in.root_position[2]};
    const Q4 root_q{in.root_position[3], in.root_position[4], in.root_position[5], in.root_position[6]};
    const V3 omega{in.root_velocity[3], in.root_velocity[4], in.root_velocity[5]};
    const V3 alpha{in.root_acceleration[3], in.root_acceleration[4], in.root_acceleration[5]};
    for (size_t k = 0; k < n_nodes; ++k) {
        const Q4 qn = tangent_twist_to_quat(tangents[k], 0.);
        V3 p{node_coords[k][0] + root_loc[0], node_coords[k][1] + root_loc[1], node_coords[k][2] + root_loc[2]};
        const V3 rr0 = rotate(root_q, {p[0] - root_loc[0], p[1] - root_loc[1], p[2] - root_loc[2]});
        p = {rr0[0] + root_loc[0], rr0[1] + root_loc[1], rr0[2] + root_loc[2]};
        const Q4 q = compose(root_q, qn);
        const V3 r{p[0] - root_loc[0], p[1] - root_loc[1], p[2] - root_loc[2]};
        const V3 wxr = cross(omega, r), axr = cross(alpha, r), wwr = cross(omega, cross(omega, r));
        const double x0[7] = {p[0], p[1], p[2], q[0], q[1], q[2], q[3]};
        const double u[7] = {0., 0., 0., 1., 0., 0., 0.};
        const double v[6] = {in.root_velocity[0] + wxr[0], in.root_velocity[1] + wxr[1], in.root_velocity[2] + wxr[2],
                             omega[0], omega[1], omega[2]};
        const double vd[6] = {in.root_acceleration[0] + axr[0] + wwr[0], in.root_acceleration[1] + axr[1] + wwr[1],
                              in.root_acceleration[2] + axr[2] + wwr[2], alpha[0], alpha[1], alpha[2]};
        const int64_t id = kyn_model_add_node(m, x0, u, v, vd, (B.node_xi[k] + 1.) / 2.);
        if (id < 0) throw BladeError(static_cast<int>(id), "kyn_model_add_node failed");
        B.node_ids.push_back(static_cast<size_t>(id));
    }

    // BuildBeamSections (beam.cpp:210-281): sections rotated by the interpolated twist (the
    // twist values are taken in degrees here), optional refinement sections
    std::vector<double> tg(in.twist_grid, in.twist_grid + in.num_twist_points), tv(in.twist, in.twist + in.num_twist_points);
    auto q_twist = [&](double loc) { return rotvec_to_quat({linear_interp(loc, tg, tv) * kPi / 180., 0., 0.}); };
    auto section = [&](size_t k, const double* base) {
        M6 mtx;
        std::copy(base + 36 * k, base + 36 * k + 36, mtx.begin());
        return z_axis ? rotate_matrix6(mtx, q_z_to_x) : mtx;
    };
    std::vector<double> s_s, s_m, s_c;
    auto push = [&](double s, const M6& M, const M6& C) {
        s_s.push_back(s);
        s_m.insert(s_m.end(), M.begin(), M.end());
        s_c.insert(s_c.end(), C.begin(), C.end());
    };
    {
        const Q4 q = q_twist(in.section_grid[0]);
        push(in.section_grid[0], rotate_matrix6(section(0, in.section_mass), q), rotate_matrix6(section(0, in.section_stiffness), q));
    }
    for (size_t k = 1; k < in.num_sections; ++k) {
        const double loc = in.section_grid[k];
        const M6 Mk = section(k, in.section_mass), Ck = section(k, in.section_stiffness);
        const M6 Ml = section(k - 1, in.section_mass), Cl = section(k - 1, in.section_stiffness);
        for (size_t r = 0; r < in.section_refinement; ++r) {
            const double a = static_cast<double>(r + 1) / static_cast<double>(in.section_refinement + 1);
            const double g = (1. - a) * in.section_grid[k - 1] + a * loc;
            M6 Mi, Ci;
            for (int c = 0; c < 36; ++c) {
                Mi[c] = (1. - a) * Ml[c] + a * Mk[c];
                Ci[c] = (1. - a) * Cl[c] + a * Ck[c];
            }
            const Q4 q = q_twist(loc);
            push(g, rotate_matrix6(Mi, q), rotate_matrix6(Ci, q));
        }
        const Q4 q = q_twist(loc);
        push(loc, rotate_matrix6(Mk, q), rotate_matrix6(Ck, q));
    }
    const auto quad = trapezoidal_quadrature(s_s);
    B.beam = kyn_model_add_beam_element(m, B.node_ids.size(), B.node_ids.data(), s_s.size(), s_s.data(), s_m.data(),
                                        s_c.data(), quad.size() / 2, quad.data());
    if (B.beam < 0) throw BladeError(static_cast<int>(B.beam), "kyn_model_add_beam_element failed");
    if (in.prescribe_root_motion) {  // SetupRootNode (beam.cpp:166-172)
        B.root_constraint = kyn_model_add_constraint(m, KYN_PRESCRIBED_BC, 0, B.node_ids.front(), nullptr, nullptr);
        if (B.root_constraint < 0) throw BladeError(static_cast<int>(B.root_constraint), "kyn_model_add_constraint failed");
    }
    return B;
}

template <typename F>
int guarded(F&& f) {
    try {
        kyn::set_last_error("");
        return f();
    } catch (const BladeError& e) {
        kyn::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        kyn::set_last_error(e.what());
        return KYN_ERR_INVALID;
    }
}

}  // namespace

struct kyn_blade_interface_s {
    kyn_system sys{};
    Built built;
    size_t n{};
    std::vector<double> loads;   // NodeData::loads [n][6]
    std::vector<double> motion;  // packed x, q, v, vd [n][26]
    ~kyn_blade_interface_s() {
        if (sys) kyn_system_destroy(sys);
    }
    void refresh() {  // HostState::CopyFromState + Beam::GetMotion
        const int rc = kyn_get_motion(sys, motion.data());
        if (rc != KYN_OK) throw BladeError(rc, kyn_last_error());
    }
};

extern "C" {

int kyn_beam_build_model(kyn_model m, const kyn_beam_input* in, int64_t* beam_element_id, int64_t* root_constraint_id) {
    return guarded([&] {
        if (!m || !in) throw BladeError(KYN_ERR_INVALID, "null argument");
        const Built b = build_beam(m, *in);
        if (beam_element_id) *beam_element_id = b.beam;
        if (root_constraint_id) *root_constraint_id = b.root_constraint;
        return KYN_OK;
    });
}

kyn_blade_interface kyn_blade_interface_create(const kyn_solution_input* solution, const kyn_beam_input* blade,
                                               const kyn_exec_options* options) {
    kyn_blade_interface out = nullptr;
    kyn_model m = nullptr;
    const int rc = guarded([&] {
        if (!solution || !blade) throw BladeError(KYN_ERR_INVALID, "null argument");
        auto bi = new kyn_blade_interface_s();
        out = bi;
        m = kyn_model_create();
        kyn_model_set_gravity(m, solution->gravity);
        bi->built = build_beam(m, *blade);
        const kyn_step_params p{solution->dynamic_solve, solution->max_iter, solution->time_step, solution->rho_inf,
                                solution->absolute_error_tolerance, solution->relative_error_tolerance};
        const kyn_exec_options def{KYN_EXEC_FUSED, KYN_TANGENT_FP64, 0, 0.};
        bi->sys = kyn_system_create(&m, 1, &p, options ? options : &def);
        if (!bi->sys) throw BladeError(KYN_ERR_INVALID, std::string("kyn_system_create failed: ") + kyn_last_error());
        bi->n = kyn_system_num_nodes(bi->sys);
        bi->loads.assign(bi->n * 6, 0.);
        bi->motion.assign(bi->n * 26, 0.);
        bi->refresh();
        return KYN_OK;
    });
    if (m) kyn_model_destroy(m);
    if (rc != KYN_OK) {
        delete out;
        return nullptr;
    }
    return out;
}

void kyn_blade_interface_destroy(kyn_blade_interface bi) { delete bi; }
size_t kyn_blade_interface_num_nodes(kyn_blade_interface bi) { return bi ? bi->n : 0; }
kyn_system kyn_blade_interface_system(kyn_blade_interface bi) { return bi ? bi->sys : nullptr; }

int kyn_blade_interface_set_node_loads(kyn_blade_interface bi, size_t node, const double loads[6]) {
    if (!bi || !loads || node >= bi->n) return KYN_ERR_INVALID;
    std::copy(loads, loads + 6, bi->loads.begin() + 6 * node);
    return KYN_OK;
}
int kyn_blade_interface_get_node_loads(kyn_blade_interface bi, size_t node, double loads[6]) {
    if (!bi || !loads || node >= bi->n) return KYN_ERR_INVALID;
    std::copy(bi->loads.begin() + 6 * node, bi->loads.begin() + 6 * node + 6, loads);
    return KYN_OK;
}
// Beam::AddPointLoad (beam.cpp:33-45): distributed with the node shape-function derivative
// weights at s, as the reference does (GetNodeWeights, beam.cpp:26-31)
int kyn_blade_interface_add_point_load(kyn_blade_interface bi, double s, const double loads[6]) {
    return guarded([&] {
        if (!bi || !loads) throw BladeError(KYN_ERR_INVALID, "null argument");
        if (s < 0. || s > 1.) throw BladeError(KYN_ERR_INVALID, "Invalid position: " + std::to_string(s));
        std::vector<double> w;
        lagrange_deriv_weights(2. * s - 1., bi->built.node_xi, w);
        for (size_t k = 0; k < bi->n; ++k)
            for (int c = 0; c < 6; ++c) bi->loads[6 * k + c] += w[k] * loads[c];
        return KYN_OK;
    });
}
int kyn_blade_interface_clear_loads(kyn_blade_interface bi) {
    if (!bi) return KYN_ERR_INVALID;
    std::fill(bi->loads.begin(), bi->loads.end(), 0.);
    return KYN_OK;
}

// BladeInterface::Step (blade_interface.cpp:54-82)
int kyn_blade_interface_step(kyn_blade_interface bi) {
    return guarded([&] {
        if (!bi) throw BladeError(KYN_ERR_INVALID, "null blade interface");
        int rc = kyn_set_nodal_loads(bi->sys, bi->loads.data());
        if (rc != KYN_OK) throw BladeError(rc, kyn_last_error());
        rc = kyn_step(bi->sys);
        if (rc < 0) throw BladeError(rc, kyn_last_error());
        if (rc != KYN_OK) return KYN_NOT_CONVERGED;
        bi->refresh();
        return KYN_OK;
    });
}

int kyn_blade_interface_get_node_motion(kyn_blade_interface bi, size_t node, double position[7], double displacement[7],
                                        double velocity[6], double acceleration[6]) {
    if (!bi || node >= bi->n) return KYN_ERR_INVALID;
    const double* o = bi->motion.data() + 26 * node;
    if (position) std::copy(o, o + 7, position);
    if (displacement) std::copy(o + 7, o + 14, displacement);
    if (velocity) std::copy(o + 14, o + 20, velocity);
    if (acceleration) std::copy(o + 20, o + 26, acceleration);
    return KYN_OK;
}

int kyn_blade_interface_save_state(kyn_blade_interface bi) {
    if (!bi) return KYN_ERR_INVALID;
    return kyn_save_state(bi->sys);
}
int kyn_blade_interface_restore_state(kyn_blade_interface bi) {
    return guarded([&] {
        if (!bi) throw BladeError(KYN_ERR_INVALID, "null blade interface");
        const int rc = kyn_restore_state(bi->sys);
        if (rc != KYN_OK) throw BladeError(rc, kyn_last_error());
        bi->refresh();
        return KYN_OK;
    });
}
int kyn_blade_interface_set_root_displacement(kyn_blade_interface bi, const double u[7]) {
    return guarded([&] {
        if (!bi || !u) throw BladeError(KYN_ERR_INVALID, "null argument");
        if (bi->built.root_constraint < 0) throw BladeError(KYN_ERR_INVALID, "prescribed root motion was not enabled");
        const int rc = kyn_set_constraint_displacement(bi->sys, static_cast<size_t>(bi->built.root_constraint), u);
        if (rc != KYN_OK) throw BladeError(rc, kyn_last_error());
        return KYN_OK;
    });
}

}  // extern "C"
