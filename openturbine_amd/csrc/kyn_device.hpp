// Device-side math of the beam Newton path (gfx950).  Each routine restates the reference
// functor named in its comment; all arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "kyn_system.hpp"

namespace kyn {

#define KYN_DEV __device__ __forceinline__

// ---- quaternion / small-matrix algebra (src/math/quaternion_operations.hpp,
//      vector_operations.hpp, matrix_operations.hpp) ----
KYN_DEV void quat_to_rot(const double* q, double* R) {  // QuaternionToRotationMatrix :19-32
    R[0] = q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
    R[1] = 2. * (q[1] * q[2] - q[0] * q[3]);
    R[2] = 2. * (q[1] * q[3] + q[0] * q[2]);
    R[3] = 2. * (q[1] * q[2] + q[0] * q[3]);
    R[4] = q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3];
    R[5] = 2. * (q[2] * q[3] - q[0] * q[1]);
    R[6] = 2. * (q[1] * q[3] - q[0] * q[2]);
    R[7] = 2. * (q[2] * q[3] + q[0] * q[1]);
    R[8] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
}
KYN_DEV void quat_rotate(const double* q, const double* v, double* r) {  // RotateVectorByQuaternion :113
    r[0] = (q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3]) * v[0] +
           2. * (q[1] * q[2] - q[0] * q[3]) * v[1] + 2. * (q[1] * q[3] + q[0] * q[2]) * v[2];
    r[1] = 2. * (q[1] * q[2] + q[0] * q[3]) * v[0] +
           (q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3]) * v[1] +
           2. * (q[2] * q[3] - q[0] * q[1]) * v[2];
    r[2] = 2. * (q[1] * q[3] - q[0] * q[2]) * v[0] + 2. * (q[2] * q[3] + q[0] * q[1]) * v[1] +
           (q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]) * v[2];
}
KYN_DEV void quat_compose(const double* a, const double* b, double* c) {  // QuaternionCompose :203
    c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    c[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    c[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    c[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}
KYN_DEV void quat_inverse(const double* q, double* r) {  // QuaternionInverse :171
    const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    r[0] = q[0] / len;
    r[1] = -q[1] / len;
    r[2] = -q[2] / len;
    r[3] = -q[3] / len;
}
KYN_DEV void rotvec_to_quat(const double* phi, double* q) {  // RotationVectorToQuaternion :230
    const double angle = sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    const double c = cos(angle / 2.);
    const double f = (fabs(angle) < 1.e-12) ? 0. : sin(angle / 2.) / angle;
    q[0] = c;
    q[1] = phi[0] * f;
    q[2] = phi[1] * f;
    q[3] = phi[2] * f;
}
KYN_DEV void tilde(const double* v, double* M) {  // VecTilde
    M[0] = 0.; M[1] = -v[2]; M[2] = v[1];
    M[3] = v[2]; M[4] = 0.; M[5] = -v[0];
    M[6] = -v[1]; M[7] = v[0]; M[8] = 0.;
}
KYN_DEV void mm3(const double* A, const double* B, double* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
KYN_DEV void mtm3(const double* A, const double* B, double* C) {  // A^T B
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
KYN_DEV void mmt3(const double* A, const double* B, double* C) {  // A B^T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
}
KYN_DEV void mv3(const double* A, const double* x, double* y) {
#pragma unroll
    for (int i = 0; i < 3; ++i) y[i] = A[i * 3] * x[0] + A[i * 3 + 1] * x[1] + A[i * 3 + 2] * x[2];
}

// RotateSectionMatrix (src/system/masses/rotate_section_matrix.hpp:17-51):
// out = blkdiag(R,R) * Cs * blkdiag(R,R)^T with R from xr; Cs read with stride 1.
KYN_DEV void rotate_section(const double* R, const double* Cs, double* out) {
    double T[36];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            T[i * 6 + j] = R[i * 3] * Cs[j] + R[i * 3 + 1] * Cs[6 + j] + R[i * 3 + 2] * Cs[12 + j];
            T[(3 + i) * 6 + j] =
                R[i * 3] * Cs[18 + j] + R[i * 3 + 1] * Cs[24 + j] + R[i * 3 + 2] * Cs[30 + j];
        }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            out[i * 6 + j] = T[i * 6] * R[j * 3] + T[i * 6 + 1] * R[j * 3 + 1] + T[i * 6 + 2] * R[j * 3 + 2];
            out[i * 6 + 3 + j] =
                T[i * 6 + 3] * R[j * 3] + T[i * 6 + 4] * R[j * 3 + 1] + T[i * 6 + 5] * R[j * 3 + 2];
        }
}

// Stored tangent rotational block (src/system/calculate_tangent_operator.hpp:26-65): the
// reference stores T^T with T = [I 0; 0 I + tmp1*psi~ + tmp2*psi~^2]; only the 3x3
// rotational block of the stored matrix is non-trivial. out[9] = (I + tmp1 psi~ + tmp2 psi~^2)^T
KYN_DEV void tangent_rot_block(double h, const double* qd, double* out) {
    const double rv[3] = {h * qd[3], h * qd[4], h * qd[5]};
    const double phi = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
    const double tmp1 = (phi > 1.e-16) ? (cos(phi) - 1.) / (phi * phi) : 0.;
    const double tmp2 = (phi > 1.e-16) ? (1. - sin(phi) / phi) / (phi * phi) : 0.;
    double m2[9], m22[9];
    tilde(rv, m2);
    mm3(m2, m2, m22);
    double T[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) T[i] = tmp2 * m22[i] + tmp1 * m2[i];
    T[0] += 1.;
    T[4] += 1.;
    T[8] += 1.;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = T[j * 3 + i];
}

// CalculateDisplacement (src/state/calculate_displacement.hpp:25-50)
KYN_DEV void calc_displacement(double h, const double* qd, const double* qp, double* q) {
    q[0] = qp[0] + h * qd[0];
    q[1] = qp[1] + h * qd[1];
    q[2] = qp[2] + h * qd[2];
    const double d[3] = {h * qd[3], h * qd[4], h * qd[5]};
    double dq[4], r[4];
    rotvec_to_quat(d, dq);
    const double qpr[4] = {qp[3], qp[4], qp[5], qp[6]};
    quat_compose(dq, qpr, r);
    q[3] = r[0];
    q[4] = r[1];
    q[5] = r[2];
    q[6] = r[3];
}

// ---------------------------------------------------------------------------------------
// QP evaluation.  Inertial half: calculate_inertial_quadrature_point_values.hpp:44-126 with
// masses/{rotate_section_matrix, calculate_mass_matrix_components, calculate_inertial_force,
// calculate_gravity_force, calculate_gyroscopic_matrix, calculate_inertia_stiffness_matrix}.
// Outputs Muu, Guu, Kuu (36 each) and Fi, Fg (6 each).
// ---------------------------------------------------------------------------------------
// Results go out as soon as they are final (keeps few 6x6 blocks live): F[12..18) = Fi,
// F[18..24) = Fg, M6[0..36) = wJ Kuu, M6[72..108) = wJ (beta' Muu + gamma' Guu).
// `r` is the (already normalized) rotation the section matrix is carried by.
template <typename TT>
KYN_DEV void inertial_terms(const double* r, const double* ud, const double* w, const double* wd,
                            const double* r0, const double* Ms, const double* g, double wJ, double bp,
                            double gp, TT* M6, double* F) {
    double Muu[36], Fi[6], Fg[6];
    double xr[4], R[9];
    quat_compose(r, r0, xr);
    quat_to_rot(xr, R);
    rotate_section(R, Ms, Muu);
    const double m = Muu[0];
    const double eta[3] = {Muu[5 * 6 + 1] / m, -Muu[5 * 6 + 0] / m, Muu[4 * 6 + 0] / m};
    double rho[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) rho[i * 3 + j] = Muu[(3 + i) * 6 + 3 + j];
    double et[9], wt[9], wdt[9];
    tilde(eta, et);
    tilde(w, wt);
    tilde(wd, wdt);
    // inertial force (calculate_inertial_force.hpp:16-53)
    double M1[9], t[9];
    mm3(wt, wt, t);
#pragma unroll
    for (int i = 0; i < 9; ++i) M1[i] = m * t[i] + m * wdt[i];
    double f1[3];
    mv3(M1, eta, f1);
#pragma unroll
    for (int i = 0; i < 3; ++i) Fi[i] = f1[i] + m * ud[i];
    const double v1[3] = {m * ud[0], m * ud[1], m * ud[2]};
    double a[3], b[3], c[3], wr[9];
    mv3(et, v1, a);
    mv3(rho, wd, b);
    mm3(wt, rho, wr);
    mv3(wr, w, c);
#pragma unroll
    for (int i = 0; i < 3; ++i) Fi[3 + i] = a[i] + b[i] + c[i];
    // gravity (calculate_gravity_force.hpp:15-27)
    const double fg[3] = {m * g[0], m * g[1], m * g[2]};
    double tg[3];
    mv3(et, fg, tg);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Fg[i] = fg[i];
        Fg[3 + i] = tg[i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        F[12 + i] = Fi[i];
        F[18 + i] = Fg[i];
    }
    // gyroscopic (calculate_gyroscopic_matrix.hpp:20-60); only the right 6x3 column is nonzero
    {
        double Guu[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) Guu[i] = 0.;
        const double me[3] = {m * eta[0], m * eta[1], m * eta[2]};
        double v2[3], A[9], B[9], tm[9];
        mv3(wt, me, v2);
        tilde(v2, A);
        tilde(me, tm);
        mmt3(wt, tm, B);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Guu[i * 6 + 3 + j] = B[i * 3 + j] + A[j * 3 + i];
        double rw[3], C[9];
        mv3(rho, w, rw);
        tilde(rw, C);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Guu[(3 + i) * 6 + 3 + j] = wr[i * 3 + j] - C[i * 3 + j];
#pragma unroll
        for (int c = 0; c < 36; ++c) M6[72 + c] = (TT)(wJ * (bp * Muu[c] + gp * Guu[c]));
    }
    // inertia stiffness (calculate_inertia_stiffness_matrix.hpp:19-60)
    {
        double N1[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) N1[i] = t[i] + wdt[i];
        const double me[3] = {m * eta[0], m * eta[1], m * eta[2]};
        double M2[9], K12[9];
        tilde(me, M2);
        mmt3(N1, M2, K12);
        double ut[9], k22[9], t1[9], rwd[3], t2[9], rw[3], t3[9], t4[9], m1[9], t5[9];
        tilde(ud, ut);
        mm3(rho, wdt, k22);
        mm3(ut, M2, t1);
#pragma unroll
        for (int i = 0; i < 9; ++i) k22[i] = t1[i] + k22[i];
        mv3(rho, wd, rwd);
        tilde(rwd, t2);
#pragma unroll
        for (int i = 0; i < 9; ++i) k22[i] += -1. * t2[i];
        mv3(rho, w, rw);
        tilde(rw, t3);
        mm3(rho, wt, t4);
#pragma unroll
        for (int i = 0; i < 9; ++i) m1[i] = t4[i] - t3[i];
        mm3(wt, m1, t5);
#pragma unroll
        for (int i = 0; i < 9; ++i) k22[i] = t5[i] + k22[i];
        double Kuu[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) Kuu[i] = 0.;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Kuu[i * 6 + 3 + j] = K12[i * 3 + j];
                Kuu[(3 + i) * 6 + 3 + j] = k22[i * 3 + j];
            }
#pragma unroll
        for (int c = 0; c < 36; ++c) M6[c] = (TT)(wJ * Kuu[c]);
    }
}

// Beam QP (src/system/beams/calculate_quadrature_point_values.hpp): the interpolated QP rotation
// is normalized before use (interpolate_QP_rotation.hpp)
template <typename TT>
KYN_DEV void qp_inertial(const double* r_in /*4, unnormalized*/, const double* ud,
                         const double* w, const double* wd, const double* r0, const double* Ms,
                         const double* g, double wJ, double bp, double gp, TT* M6, double* F) {
    double r[4];
    const double rl = sqrt(r_in[0] * r_in[0] + r_in[1] * r_in[1] + r_in[2] * r_in[2] + r_in[3] * r_in[3]);
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = r_in[c] / rl;
    inertial_terms<TT>(r, ud, w, wd, r0, Ms, g, wJ, bp, gp, M6, F);
}

// Point mass (src/system/masses/calculate_quadrature_point_values.hpp:41-152): the node's own
// rotation q[3..7) composed with its reference rotation, no normalization; residual Fi - Fg,
// matrix beta' Muu + gamma' Guu + Kuu T with T = blkdiag(I, trr) the node's stored tangent
// (Gemm(1, Kuu, T, 1, STpI): the product is formed first, then added)
KYN_DEV void mass_terms(const double* r, const double* ud, const double* w, const double* wd,
                        const double* r0, const double* Ms, const double* g, double bp, double gp,
                        const double* trr, double* res, double* mat) {
    double M6[108], F[24];
    inertial_terms<double>(r, ud, w, wd, r0, Ms, g, 1., bp, gp, M6, F);
#pragma unroll
    for (int c = 0; c < 6; ++c) res[c] = F[12 + c] - F[18 + c];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) mat[i * 6 + j] = M6[72 + i * 6 + j] + M6[i * 6 + j];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double kt = M6[i * 6 + 3] * trr[j] + M6[i * 6 + 4] * trr[3 + j] + M6[i * 6 + 5] * trr[6 + j];
            mat[i * 6 + 3 + j] = M6[72 + i * 6 + 3 + j] + kt;
        }
    }
}

// Spring (src/system/springs/calculate_quadrature_point_values.hpp:28-90 with
// calculate_distance_components.hpp, calculate_length.hpp, calculate_force_coefficients.hpp,
// calculate_force_vectors.hpp, calculate_stiffness_matrix.hpp): r = x0 + u2 - u1, l = |r|,
// c1 = k (l_ref / l - 1), c2 = k l_ref / l^3, f = c1 r, a = (c1 - c2 l^2) I - c2 r~ r~;
// residual [f, -f]; the 2 x 2 node blocks are [a, -a; -a, a]
KYN_DEV void spring_terms(const double* x0, const double* u1, const double* u2, double l_ref, double k,
                          double* res, double* a) {
    double r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = x0[c] + u2[c] - u1[c];
    const double l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double c1 = k * (l_ref / l - 1.);
    const double c2 = k * l_ref / (l * l * l);
    double rt[9], rr[9];
    tilde(r, rt);
    mm3(rt, rt, rr);
    const double diag = c1 - c2 * l * l;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a[i * 3 + j] = -c2 * rr[i * 3 + j] + (i == j ? diag : 0.);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        res[c] = c1 * r[c];
        res[3 + c] = -(c1 * r[c]);
    }
}

// Stiffness half: calculate_stiffness_quadrature_point_values.hpp:47-128 with
// beams/{calculate_strain, calculate_temporary_variables, calculate_force_FC,
// calculate_force_FD, calculate_Ouu, calculate_Puu, calculate_Quu}.
// Results go out as soon as they are final: F[0..6) = Fc, F[6..12) = Fd,
// M6[180..216) = (w/J) Cuu, M6[144..180) = w Ouu, M6[108..144) = w Puu, M6[36..72) = wJ Quu.
template <typename TT>
KYN_DEV void qp_stiffness(const double* r_in, const double* rp, const double* up,
                          const double* r0, const double* x0p, const double* Cs, double wJ, double w,
                          double wdJ, TT* M6, double* F) {
    double Cuu[36], Fc[6], Fd[6];
    double r[4];
    const double rl = sqrt(r_in[0] * r_in[0] + r_in[1] * r_in[1] + r_in[2] * r_in[2] + r_in[3] * r_in[3]);
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = r_in[c] / rl;
    double xr[4], R[9];
    quat_compose(r, r0, xr);
    quat_to_rot(xr, R);
    rotate_section(R, Cs, Cuu);
    // strain (calculate_strain.hpp:17-47)
    double Rx[3];
    quat_rotate(r, x0p, Rx);
#pragma unroll
    for (int i = 0; i < 3; ++i) Rx[i] += -1. * up[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) Rx[i] += -1. * x0p[i];
    const double E[12] = {-r[1], r[0], -r[3], r[2], -r[2], r[3], r[0], -r[1], -r[3], -r[2], r[1], r[0]};
    double strain[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        strain[i] = -Rx[i];
        strain[3 + i] = 2. * (E[i * 4] * rp[0] + E[i * 4 + 1] * rp[1] + E[i * 4 + 2] * rp[2] + E[i * 4 + 3] * rp[3]);
    }
    const double x0pup[3] = {x0p[0] + up[0], x0p[1] + up[1], x0p[2] + up[2]};
    double xs[9];
    tilde(x0pup, xs);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = 0.;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += Cuu[i * 6 + k] * strain[k];
        Fc[i] = s;
    }
    Fd[0] = 0.;
    Fd[1] = 0.;
    Fd[2] = 0.;
#pragma unroll
    for (int i = 0; i < 3; ++i) Fd[3 + i] = xs[i] * Fc[0] + xs[3 + i] * Fc[1] + xs[6 + i] * Fc[2];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        F[i] = Fc[i];
        F[6 + i] = Fd[i];
    }
    double Nt[9], Mt[9];
    tilde(Fc, Nt);
    tilde(Fc + 3, Mt);
    double C11[9], C12[9], C21[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            C11[i * 3 + j] = Cuu[i * 6 + j];
            C12[i * 3 + j] = Cuu[i * 6 + 3 + j];
            C21[i * 3 + j] = Cuu[(3 + i) * 6 + j];
        }
#pragma unroll
    for (int c = 0; c < 36; ++c) M6[180 + c] = (TT)(wdJ * Cuu[c]);
    double Ouu[36], Puu[36], Quu[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) {
        Ouu[i] = 0.;
        Puu[i] = 0.;
        Quu[i] = 0.;
    }
    double a[9], b[9];
    mm3(C11, xs, a);
    mm3(C21, xs, b);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Ouu[i * 6 + 3 + j] = a[i * 3 + j] - Nt[i * 3 + j];
            Ouu[(3 + i) * 6 + 3 + j] = b[i * 3 + j] - Mt[i * 3 + j];
        }
#pragma unroll
    for (int c = 0; c < 36; ++c) M6[144 + c] = (TT)(w * Ouu[c]);
    double c[9], d[9];
    mtm3(xs, C11, c);
    mtm3(xs, C12, d);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Puu[(3 + i) * 6 + j] = c[i * 3 + j] + Nt[i * 3 + j];
            Puu[(3 + i) * 6 + 3 + j] = d[i * 3 + j];
        }
#pragma unroll
    for (int k = 0; k < 36; ++k) M6[108 + k] = (TT)(w * Puu[k]);
    double m1[9], q22[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) m1[i] = a[i] - Nt[i];
    mtm3(xs, m1, q22);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Quu[(3 + i) * 6 + 3 + j] = q22[i * 3 + j];
#pragma unroll
    for (int k = 0; k < 36; ++k) M6[36 + k] = (TT)(wJ * Quu[k]);
}

// AX_Matrix / AxialVectorOfMatrix (src/math/matrix_operations.hpp:18-49)
KYN_DEV void ax_matrix(const double* A, double* R) {
    const double trace = (A[0] + A[4] + A[8]) / 2.;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = -A[i * 3 + j] / 2.;
        R[i * 3 + i] += trace;
    }
}

// Constraint residual and gradients for FixedBC, PrescribedBC, RigidJoint, RevoluteJoint,
// RotationControl and the three-dof variants (src/constraints/
// calculate_constraint_residual_gradient.hpp:61-565 and the kernels
// calculate_{fixed_bc,prescribed_bc,rigid_joint,revolute_joint,rotation_control}_constraint.hpp,
// calculate_{fixed_bc,prescribed_bc,rigid_joint}_3DOF_constraint.hpp,
// calculate_revolute_joint_force.hpp).  Outputs: res[6]; tg, bg: gradient (not yet times
// tangent); sys_res[3..5] (revolute only).
KYN_DEV void constraint_eval(int type, const double* X0, const double* axes, const double* input,
                             const double* bq /*7*/, const double* tq /*7*/, double* res,
                             double* bg, double* tg, double* sys_res) {
#pragma unroll
    for (int i = 0; i < 36; ++i) {
        bg[i] = 0.;
        tg[i] = 0.;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) res[i] = 0.;
    double u1[3] = {0., 0., 0.}, R1[4] = {1., 0., 0., 0.};
    if (con_prescribed(type)) {
        u1[0] = input[0]; u1[1] = input[1]; u1[2] = input[2];
        R1[0] = input[3]; R1[1] = input[4]; R1[2] = input[5]; R1[3] = input[6];
    } else if (con_two_node(type)) {
        u1[0] = bq[0]; u1[1] = bq[1]; u1[2] = bq[2];
        R1[0] = bq[3]; R1[1] = bq[4]; R1[2] = bq[5]; R1[3] = bq[6];
    }
    const double u2[3] = {tq[0], tq[1], tq[2]};
    const double R2[4] = {tq[3], tq[4], tq[5], tq[6]};
    double R1t[4], R1X0[3];
    quat_inverse(R1, R1t);
    quat_rotate(R1, X0, R1X0);
#pragma unroll
    for (int k = 0; k < 3; ++k) res[k] = u2[k] + X0[k] - u1[k] - R1X0[k];
    if (con_three_dof(type)) {
        // calculate_{fixed_bc,prescribed_bc,rigid_joint}_3DOF_constraint.hpp: translational rows
        // only; target gradient I, base gradient [-I tilde(R1 X0)]
        tg[0] = 1.;
        tg[7] = 1.;
        tg[14] = 1.;
        if (type == kConRigidJoint6DOFsTo3DOFs) {
            bg[0] = -1.;
            bg[7] = -1.;
            bg[14] = -1.;
            double Xt[9];
            tilde(R1X0, Xt);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) bg[i * 6 + 3 + j] = Xt[i * 3 + j];
        }
    } else if (type != kConRevoluteJoint) {
        double R2R1t[4], C[9];
        if (type == kConRotationControl) {
            // calculate_rotation_control_constraint.hpp:85-100: C = R(R2 inv(Rc) inv(R1)) with
            // Rc the commanded rotation input(0) about the unit axis
            const double rv[3] = {axes[0] * input[0], axes[1] * input[0], axes[2] * input[0]};
            double Rc[4], RcT[4], R2RcT[4];
            rotvec_to_quat(rv, Rc);
            quat_inverse(Rc, RcT);
            quat_compose(R2, RcT, R2RcT);
            quat_compose(R2RcT, R1t, R2R1t);
        } else {
            quat_compose(R2, R1t, R2R1t);
        }
        quat_to_rot(R2R1t, C);
        res[3] = (C[7] - C[5]) / 2.;
        res[4] = (C[2] - C[6]) / 2.;
        res[5] = (C[3] - C[1]) / 2.;
        tg[0] = 1.;
        tg[7] = 1.;
        tg[14] = 1.;
        double A[9];
        ax_matrix(C, A);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tg[(3 + i) * 6 + 3 + j] = A[j * 3 + i];
        if (type == kConRigidJoint || type == kConRotationControl) {
            bg[0] = -1.;
            bg[7] = -1.;
            bg[14] = -1.;
            double Xt[9];
            tilde(R1X0, Xt);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    bg[i * 6 + 3 + j] = Xt[i * 3 + j];
                    bg[(3 + i) * 6 + 3 + j] = -A[i * 3 + j];
                }
        }
    } else {
        const double x0v[3] = {axes[0], axes[1], axes[2]};
        const double y0v[3] = {axes[3], axes[4], axes[5]};
        const double z0v[3] = {axes[6], axes[7], axes[8]};
        double x[3], y[3], z[3];
        quat_rotate(R1, x0v, x);
        quat_rotate(R2, y0v, y);
        quat_rotate(R2, z0v, z);
        res[3] = z[0] * x[0] + z[1] * x[1] + z[2] * x[2];
        res[4] = y[0] * x[0] + y[1] * x[1] + y[2] * x[2];
        tg[0] = 1.;
        tg[7] = 1.;
        tg[14] = 1.;
        const double xcz[3] = {x[1] * z[2] - x[2] * z[1], x[2] * z[0] - x[0] * z[2], x[0] * z[1] - x[1] * z[0]};
        const double xcy[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tg[3 * 6 + 3 + k] = -xcz[k];
            tg[4 * 6 + 3 + k] = -xcy[k];
        }
        bg[0] = -1.;
        bg[7] = -1.;
        bg[14] = -1.;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bg[3 * 6 + 3 + k] = xcz[k];
            bg[4 * 6 + 3 + k] = xcy[k];
        }
        double R2X0[3];
        quat_rotate(R2, x0v, R2X0);
        sys_res[3] = R2X0[0] * input[0];
        sys_res[4] = R2X0[1] * input[0];
        sys_res[5] = R2X0[2] * input[0];
    }
}

}  // namespace kyn
