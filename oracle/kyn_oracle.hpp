// ORACLE — CPU FP64 restatement of Kynema's (openturbine's) per-Newton-iteration beam path.
//
// This is TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker (or the timed CPU baseline).
// The product (openturbine_amd/) never links or calls it.
//
// Every function cites the reference file:line it restates. Formulas and operation
// order follow the reference; storage is plain row-major std::vector<double>.
//
// Pinning (see tests/test_oracle_golden.py):
//   * end-to-end Step vectors of tests/regression_tests/regression/test_solver.cpp
//     (assembled b, solution x, convergence_err[0], q_delta/v/vd/q/lambda),
//     test_cantilever_beam.cpp (tip q, 3 steps), test_rotor.cpp (IEA-15 rotor q, 1e-13);
//   * functor known answers of tests/unit_tests/system/{beams,masses}, state, math;
//   * closed forms: Gauss-Legendre exactness, Lagrange partition of unity, quaternion
//     exp/log round trip, gen-alpha coefficients, element tangent vs central finite
//     difference of the element residual, Timoshenko static tip deflection.
// Parity unpinned: none of the functions below is left without one of these checks,
// except the Gauss-Legendre generator's initial guess (the rule itself is pinned).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace kyn_oracle {

using Vec3 = std::array<double, 3>;
using Quat = std::array<double, 4>;
using Mat3 = std::array<double, 9>;   // row-major 3x3
using Mat6 = std::array<double, 36>;  // row-major 6x6

// ---------------------------------------------------------------------------------------
// Math (src/math/quaternion_operations.hpp, vector_operations.hpp, matrix_operations.hpp,
// interpolation.hpp)
// ---------------------------------------------------------------------------------------
Mat3 QuaternionToRotationMatrix(const Quat& q);
Vec3 RotateVectorByQuaternion(const Quat& q, const Vec3& v);
std::array<double, 12> QuaternionDerivative(const Quat& q);
Quat QuaternionInverse(const Quat& q);
Quat QuaternionCompose(const Quat& q1, const Quat& q2);
Quat RotationVectorToQuaternion(const Vec3& phi);  // device form (quaternion_operations.hpp:230)
Quat RotationVectorToQuaternionHost(const Vec3& phi);  // std::array form (:266)
Vec3 QuaternionToRotationVector(const Quat& q);
Mat3 VecTilde(const Vec3& v);
Mat3 AX_Matrix(const Mat3& A);
Vec3 AxialVectorOfMatrix(const Mat3& m);
void LagrangePolynomialInterpWeights(double x, const std::vector<double>& xs, std::vector<double>& w);
void LagrangePolynomialDerivWeights(double x, const std::vector<double>& xs, std::vector<double>& w);
void LinearInterpWeights(double x, const std::vector<double>& xs, std::vector<double>& w);
double LegendrePolynomial(size_t n, double x);
std::vector<double> GenerateGLLPoints(size_t order);
// Gauss-Legendre rule on [-1,1]; no reference counterpart (tests hard-code rules).
void GaussLegendreRule(size_t n, std::vector<double>& xi, std::vector<double>& w);

// ---------------------------------------------------------------------------------------
// Beam QP functors (src/system/beams/*.hpp, src/system/masses/*.hpp)
// ---------------------------------------------------------------------------------------
Mat6 RotateSectionMatrix(const Quat& xr, const Mat6& Cstar);
std::array<double, 6> CalculateStrain(const Vec3& x0p, const Vec3& up, const Quat& r, const Quat& rp);
Mat3 CalculateTemporaryVariables(const Vec3& x0p, const Vec3& up);
std::array<double, 6> CalculateForceFC(const Mat6& Cuu, const std::array<double, 6>& strain);
std::array<double, 6> CalculateForceFD(const Mat3& x0pupSS, const std::array<double, 6>& FC);
Mat6 CalculateOuu(const Mat6& Cuu, const Mat3& x0pupSS, const Mat3& M_tilde, const Mat3& N_tilde);
Mat6 CalculatePuu(const Mat6& Cuu, const Mat3& x0pupSS, const Mat3& N_tilde);
Mat6 CalculateQuu(const Mat6& Cuu, const Mat3& x0pupSS, const Mat3& N_tilde);
Vec3 CalculateEta(const Mat6& Muu);
Mat3 CalculateRho(const Mat6& Muu);
std::array<double, 6> CalculateInertialForce(double m, const Vec3& u_ddot, const Vec3& omega,
                                             const Vec3& omega_dot, const Vec3& eta,
                                             const Mat3& eta_tilde, const Mat3& rho,
                                             const Mat3& omega_tilde, const Mat3& omega_dot_tilde);
std::array<double, 6> CalculateGravityForce(double m, const Vec3& g, const Mat3& eta_tilde);
Mat6 CalculateGyroscopicMatrix(double m, const Vec3& omega, const Vec3& eta, const Mat3& rho,
                               const Mat3& omega_tilde);
Mat6 CalculateInertiaStiffnessMatrix(double m, const Vec3& u_ddot, const Vec3& omega,
                                     const Vec3& omega_dot, const Vec3& eta, const Mat3& rho,
                                     const Mat3& omega_tilde, const Mat3& omega_dot_tilde);
Mat6 CalculateTangentOperatorT(double h, const double* q_delta6);  // stores T^T
void IntegrateResidualVector(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                             const double* dphi, const double* Fc, const double* Fd, const double* Fi,
                             const double* Fe, const double* Fg, const double* node_FX, double* residual);
void IntegrateStiffnessMatrix(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                              const double* dphi, const Mat6* Kuu, const Mat6* Puu, const Mat6* Cuu, const Mat6* Ouu,
                              const Mat6* Quu, double* out);
void IntegrateInertiaMatrix(size_t N, size_t Q, const double* weight, const double* jacobian, const double* phi,
                            double beta_prime, double gamma_prime, const Mat6* Muu, const Mat6* Guu, double* out);
void CalculateSystemMatrix(size_t N, const double* M, const double* K, const double* const* T, double* S);
void CalculateNextStateNode(double h, double alpha_f, double alpha_m, double beta, double gamma, double* qd,
                            double* v, double* vd, double* a);
void CalculateDisplacementNode(double h, const double* qd, const double* qp, double* q);
void UpdateDynamicPredictionNode(double h, double beta_prime, double gamma_prime, const double* x, size_t nd,
                                 double* qd, double* v, double* vd);
void UpdateStaticPredictionNode(double h, const double* x, size_t nd, double* qd);
void UpdateAlgorithmicAccelerationNode(double alpha_f, double alpha_m, const double* vd, double* a);
void UpdateGlobalPositionNode(const double* q, const double* x0, double* x);
void CalculateJacobian(size_t N, size_t Q, const double* node_x0, const double* dphi, double* J, double* x0p);

// ---------------------------------------------------------------------------------------
// Model inputs (src/model/model.hpp, node.hpp; src/elements/beams/beam_element.hpp)
// ---------------------------------------------------------------------------------------
enum class CType : int {  // src/constraints/constraint_type.hpp:9-21
    None = 0,
    FixedBC = 1,
    PrescribedBC = 2,
    RigidJoint = 3,
    RevoluteJoint = 4,
    RotationControl = 5,
    FixedBC3DOFs = 6,
    PrescribedBC3DOFs = 7,
    RigidJoint6DOFsTo3DOFs = 8
};
// constraints between a base and a target node (constraint_type.hpp:28-33)
inline bool TwoNode(CType t) {
    return t == CType::RigidJoint || t == CType::RevoluteJoint || t == CType::RotationControl ||
           t == CType::RigidJoint6DOFsTo3DOFs;
}
// constraints on a target node with three (translational) dofs (constraint_type.hpp:64-70)
inline bool ThreeDof(CType t) {
    return t == CType::FixedBC3DOFs || t == CType::PrescribedBC3DOFs || t == CType::RigidJoint6DOFsTo3DOFs;
}
inline bool Prescribed(CType t) { return t == CType::PrescribedBC || t == CType::PrescribedBC3DOFs; }

struct NodeIn {
    std::array<double, 7> x0{0., 0., 0., 1., 0., 0., 0.};
    std::array<double, 7> u{0., 0., 0., 1., 0., 0., 0.};
    std::array<double, 6> v{};
    std::array<double, 6> vd{};
    double s{0.};
};
struct SectionIn {
    double s;
    Mat6 M;
    Mat6 C;
};
struct BeamIn {
    std::vector<size_t> nodes;
    std::vector<SectionIn> sections;
    std::vector<std::array<double, 2>> quad;
};
struct ConstraintIn {
    CType type;
    size_t base;
    size_t target;
    Vec3 axis{0., 0., 0.};
    std::array<double, 7> init{0., 0., 0., 1., 0., 0., 0.};
};
struct MassIn {  // src/elements/masses/mass_element.hpp
    size_t node;
    Mat6 M;
};
struct SpringIn {  // src/elements/springs/spring_element.hpp
    size_t n1, n2;
    double k, l0;
};
struct ModelIn {
    std::vector<NodeIn> nodes;
    std::vector<BeamIn> beams;
    std::vector<MassIn> masses;
    std::vector<SpringIn> springs;
    std::vector<ConstraintIn> constraints;
    Vec3 gravity{0., 0., 0.};
};

// src/step/step_parameters.hpp:37-52
struct StepParameters {
    bool is_dynamic_solve;
    size_t max_iter;
    double h, alpha_m, alpha_f, gamma, beta, gamma_prime, beta_prime, conditioner;
    double absolute_convergence_tol, relative_convergence_tol;
    StepParameters(bool dyn, size_t max_iter_, double h_, double rho_inf, double a_tol = 1e-5,
                   double r_tol = 1e-3);
};

// ---------------------------------------------------------------------------------------
// System (State + Beams + Constraints + Solver of the reference, host-resident)
// ---------------------------------------------------------------------------------------
struct BeamElem {
    size_t N{}, Q{};
    std::vector<size_t> node_idx;                 // [N]
    std::vector<double> weight, jacobian;         // [Q]
    std::vector<double> shape_interp, shape_deriv;  // [N][Q]
    std::vector<double> qp_x0, qp_x0_prime;       // [Q][3]
    std::vector<double> qp_r0;                    // [Q][4]
    std::vector<double> qp_Mstar, qp_Cstar;       // [Q][36]
    std::vector<double> qp_Fe;                    // [Q][6]
    std::vector<double> node_FX;                  // [N][6]
    std::vector<double> residual;                 // [N][6]
    std::vector<double> sysmat;                   // [N][N][36]
};

struct MassElem {  // src/elements/masses/masses.hpp
    size_t node{};
    Mat6 Mstar{};
    std::array<double, 7> x0{};
    std::array<double, 6> residual{};
    Mat6 sysmat{};
};
struct SpringElem {  // src/elements/springs/springs.hpp
    size_t n1{}, n2{};
    Vec3 x0{};  // node 2 x0 - node 1 x0
    double l_ref{}, k{};
    std::array<double, 6> residual{};  // [2][3]
    std::array<double, 36> stiff{};    // [2][2][3][3]
};

struct Constraint {
    CType type{};
    size_t base{}, target{};
    Vec3 X0{};
    Mat3 axes{};
    std::array<double, 7> input{0., 0., 0., 1., 0., 0., 0.};
    double control{0.};  // revolute torque / commanded rotation (input(0) at UpdateViews)
    std::array<double, 3> output{};
    std::array<double, 6> lambda{}, res{}, b_lres{}, t_lres{}, sys_res{};
    Mat6 b_grad{}, t_grad{}, b_grad_T{}, t_grad_T{};
    size_t row_first{}, row_count{};
    int base_sig{}, target_sig{};
    size_t base_dof{}, target_dof{};
};

struct Block {                     // one independent connected component of the KKT system
    std::vector<size_t> gdof;     // local index -> global dof (system dofs, then constraint rows)
    size_t n{};
    std::vector<double> A;        // n x n row-major
    std::vector<int> piv;
};

struct System {
    size_t num_nodes{};
    size_t time_step{};
    std::vector<double> x0, x, q_prev, q;        // [n][7]
    std::vector<double> q_delta, v, vd, a, f;    // [n][6]
    std::vector<double> tangent;                 // [n][36]
    std::vector<int> sig;                        // freedom signature bits
    std::vector<size_t> active_dofs, dof_map;
    Vec3 gravity{};
    std::vector<BeamElem> beams;
    std::vector<MassElem> masses;
    std::vector<SpringElem> springs;
    std::vector<Constraint> cons;
    size_t num_system_dofs{}, num_dofs{};
    std::vector<double> b, xsol;                 // [num_dofs]
    std::vector<Block> blocks;
    std::vector<size_t> gdof_block, gdof_local;
    std::vector<double> convergence_err;
};

System CreateSystem(const ModelIn& model);

// CRS graph of the full matrix (src/solver/create_full_matrix.hpp:19-74): the inputs the reference
// passes to ComputeRowPtrs / ComputeColInds, per node, element (beams, then masses, then springs,
// as in Elements) and constraint
struct CrsGraphIn {
    size_t num_system_dofs{}, num_dofs{};
    std::vector<size_t> active_dofs, node_freedom_map_table;
    std::vector<std::vector<size_t>> element_nodes;
    std::vector<size_t> base_active_dofs, target_active_dofs;
    std::vector<std::array<size_t, 6>> base_node_freedom_table, target_node_freedom_table;
    std::vector<std::pair<size_t, size_t>> row_range;
};
CrsGraphIn CrsGraphInput(const System& s);
std::vector<size_t> ComputeRowPtrs(const CrsGraphIn& g);
std::vector<size_t> ComputeColInds(const CrsGraphIn& g, const std::vector<size_t>& row_ptrs);
void SortCrsGraph(const std::vector<size_t>& row_ptrs, std::vector<size_t>& col_inds);

// Step pieces (src/step/*.hpp)
void PredictNextState(const StepParameters& p, System& s);
void ResetConstraints(System& s);
void ResetSolver(System& s);
void UpdateTangentOperator(const StepParameters& p, System& s);
void UpdateSystemVariablesBeams(const StepParameters& p, System& s);
void UpdateSystemVariablesMasses(const StepParameters& p, System& s);
void UpdateSystemVariablesSprings(System& s);
void AssembleSystemResidual(System& s);
void AssembleSystemMatrix(const StepParameters& p, System& s);
void UpdateConstraintVariables(System& s);
void AssembleConstraintsMatrix(System& s);
void AssembleConstraintsResidual(System& s);
void SolveSystem(const StepParameters& p, System& s);
double CalculateConvergenceError(const StepParameters& p, const System& s);
void UpdateStatePrediction(const StepParameters& p, System& s);
void UpdateConstraintPrediction(System& s);
void StepEpilogue(const StepParameters& p, System& s);
bool Step(const StepParameters& p, System& s);

// Element-level evaluation for one beam element: fills residual and sysmat
// (src/system/beams/calculate_quadrature_point_values.hpp:50-172)
void CalculateQuadraturePointValues(double beta_prime, double gamma_prime, const System& s,
                                    BeamElem& e);

// Point mass and spring element evaluation (src/system/masses/calculate_quadrature_point_values.hpp,
// src/system/springs/calculate_quadrature_point_values.hpp)
void CalculateMassValues(double beta_prime, double gamma_prime, const System& s, MassElem& e);
void CalculateSpringValues(const System& s, SpringElem& e);

// Dense LU with partial pivoting (replaces the reference's dss:: sparse direct solve;
// KLU's pivot order is not pinned by the reference, any correct pivoted LU is admissible)
bool LUFactor(size_t n, double* A, int* piv);
void LUSolve(size_t n, const double* A, const int* piv, double* b);

}  // namespace kyn_oracle
