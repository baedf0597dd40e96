// ORACLE C-ABI — test infrastructure only (loaded by tests/ and bench.py's cpu_baseline leg).
// Mirrors the product's kyn_* entry points one for one (prefix orc_) so that the same
// workload builder can drive either library.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kyn_oracle.hpp"

using namespace kyn_oracle;

namespace {
thread_local std::string g_err;
struct OrcSystem {
    System sys;
    StepParameters params;
    OrcSystem(System s, StepParameters p) : sys(std::move(s)), params(p) {}
};
}  // namespace

namespace {
Mat6 m6(const double* a) {
    Mat6 m{};
    std::copy(a, a + 36, m.begin());
    return m;
}
Mat3 m3(const double* a) {
    Mat3 m{};
    std::copy(a, a + 9, m.begin());
    return m;
}
Vec3 v3(const double* a) { return {a[0], a[1], a[2]}; }
template <typename A>
void put(const A& a, double* out) {
    std::copy(a.begin(), a.end(), out);
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_model_create() { return new ModelIn(); }
void orc_model_destroy(void* m) { delete static_cast<ModelIn*>(m); }
int orc_model_set_gravity(void* m, const double* g) {
    static_cast<ModelIn*>(m)->gravity = {g[0], g[1], g[2]};
    return 0;
}
long orc_model_add_node(void* m, const double* x0, const double* u, const double* v,
                        const double* vd, double s) {
    auto* mo = static_cast<ModelIn*>(m);
    NodeIn n;
    std::copy(x0, x0 + 7, n.x0.begin());
    std::copy(u, u + 7, n.u.begin());
    std::copy(v, v + 6, n.v.begin());
    std::copy(vd, vd + 6, n.vd.begin());
    n.s = s;
    mo->nodes.push_back(n);
    return static_cast<long>(mo->nodes.size() - 1);
}
long orc_model_add_beam_element(void* m, size_t nn, const size_t* nodes, size_t ns,
                                const double* sec_s, const double* sec_M, const double* sec_C,
                                size_t nq, const double* qp) {
    auto* mo = static_cast<ModelIn*>(m);
    BeamIn b;
    b.nodes.assign(nodes, nodes + nn);
    for (size_t k = 0; k < ns; ++k) {
        SectionIn sec{};
        sec.s = sec_s[k];
        std::copy(sec_M + k * 36, sec_M + k * 36 + 36, sec.M.begin());
        std::copy(sec_C + k * 36, sec_C + k * 36 + 36, sec.C.begin());
        b.sections.push_back(sec);
    }
    for (size_t q = 0; q < nq; ++q) b.quad.push_back({qp[2 * q], qp[2 * q + 1]});
    mo->beams.push_back(std::move(b));
    return static_cast<long>(mo->beams.size() - 1);
}
long orc_model_add_constraint(void* m, int type, size_t base, size_t target, const double* axis,
                              const double* init) {
    auto* mo = static_cast<ModelIn*>(m);
    ConstraintIn c{static_cast<CType>(type), base, target};
    if (axis) std::copy(axis, axis + 3, c.axis.begin());
    if (init) std::copy(init, init + 7, c.init.begin());
    mo->constraints.push_back(c);
    return static_cast<long>(mo->constraints.size() - 1);
}

long orc_model_add_mass_element(void* m, size_t node, const double* M36) {
    auto* mo = static_cast<ModelIn*>(m);
    if (node >= mo->nodes.size()) return -1;
    MassIn e{node, {}};
    std::copy(M36, M36 + 36, e.M.begin());
    mo->masses.push_back(e);
    return static_cast<long>(mo->masses.size() - 1);
}
long orc_model_add_spring_element(void* m, size_t n1, size_t n2, double k, double l0) {
    auto* mo = static_cast<ModelIn*>(m);
    if (n1 >= mo->nodes.size() || n2 >= mo->nodes.size()) return -1;
    mo->springs.push_back(SpringIn{n1, n2, k, l0});
    return static_cast<long>(mo->springs.size() - 1);
}

void* orc_system_create(void* m, int is_dynamic, size_t max_iter, double h, double rho_inf,
                        double atol, double rtol) {
    try {
        auto sys = CreateSystem(*static_cast<ModelIn*>(m));
        return new OrcSystem(std::move(sys), StepParameters(is_dynamic != 0, max_iter, h, rho_inf, atol, rtol));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_system_destroy(void* s) { delete static_cast<OrcSystem*>(s); }

int orc_step(void* s) {
    auto* o = static_cast<OrcSystem*>(s);
    return Step(o->params, o->sys) ? 0 : 1;
}
// Steps several independent systems in parallel (OpenMP over systems).
void orc_step_many(void** s, size_t n, int* status) {
#pragma omp parallel for schedule(dynamic)
    for (long i = 0; i < static_cast<long>(n); ++i) {
        auto* o = static_cast<OrcSystem*>(s[i]);
        status[i] = Step(o->params, o->sys) ? 0 : 1;
    }
}

// Runs the first assembly of a step without solving (the sequence of
// tests/regression_tests/regression/test_solver.cpp:97-112).
void orc_debug_assemble(void* s) {
    auto* o = static_cast<OrcSystem*>(s);
    PredictNextState(o->params, o->sys);
    ResetConstraints(o->sys);
    ResetSolver(o->sys);
    UpdateTangentOperator(o->params, o->sys);
    UpdateSystemVariablesBeams(o->params, o->sys);
    UpdateSystemVariablesMasses(o->params, o->sys);
    UpdateSystemVariablesSprings(o->sys);
    AssembleSystemMatrix(o->params, o->sys);
    AssembleSystemResidual(o->sys);
    UpdateConstraintVariables(o->sys);
    AssembleConstraintsMatrix(o->sys);
    AssembleConstraintsResidual(o->sys);
}

size_t orc_num_nodes(void* s) { return static_cast<OrcSystem*>(s)->sys.num_nodes; }
size_t orc_num_dofs(void* s) { return static_cast<OrcSystem*>(s)->sys.num_dofs; }
size_t orc_num_system_dofs(void* s) { return static_cast<OrcSystem*>(s)->sys.num_system_dofs; }
size_t orc_num_blocks(void* s) { return static_cast<OrcSystem*>(s)->sys.blocks.size(); }

// field: 0 x0, 1 x, 2 q_prev, 3 q (x7); 4 q_delta, 5 v, 6 vd, 7 a, 8 f (x6); 9 tangent (x36)
static std::vector<double>* field_ptr(System& s, int field, int& width) {
    switch (field) {
        case 0: width = 7; return &s.x0;
        case 1: width = 7; return &s.x;
        case 2: width = 7; return &s.q_prev;
        case 3: width = 7; return &s.q;
        case 4: width = 6; return &s.q_delta;
        case 5: width = 6; return &s.v;
        case 6: width = 6; return &s.vd;
        case 7: width = 6; return &s.a;
        case 8: width = 6; return &s.f;
        case 9: width = 36; return &s.tangent;
        default: return nullptr;
    }
}
int orc_get_state(void* s, int field, double* out) {
    int w = 0;
    auto* v = field_ptr(static_cast<OrcSystem*>(s)->sys, field, w);
    if (!v) return -1;
    std::copy(v->begin(), v->end(), out);
    return 0;
}
int orc_set_state(void* s, int field, const double* in) {
    int w = 0;
    auto* v = field_ptr(static_cast<OrcSystem*>(s)->sys, field, w);
    if (!v) return -1;
    std::copy(in, in + v->size(), v->begin());
    return 0;
}
size_t orc_time_step(void* s) { return static_cast<OrcSystem*>(s)->sys.time_step; }
int orc_set_nodal_loads(void* s, const double* f) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    std::copy(f, f + sy.num_nodes * 6, sy.f.begin());
    return 0;
}
int orc_set_element_node_loads(void* s, size_t elem, const double* fx) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    if (elem >= sy.beams.size()) return -1;
    auto& e = sy.beams[elem];
    std::copy(fx, fx + e.N * 6, e.node_FX.begin());
    return 0;
}
int orc_set_constraint_displacement(void* s, size_t id, const double* d) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    if (id >= sy.cons.size()) return -1;
    std::copy(d, d + 7, sy.cons[id].input.begin());
    return 0;
}
int orc_set_control(void* s, size_t id, double v) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    if (id >= sy.cons.size()) return -1;
    sy.cons[id].control = v;
    return 0;
}
size_t orc_get_convergence(void* s, double* out, size_t cap) {
    auto& e = static_cast<OrcSystem*>(s)->sys.convergence_err;
    const size_t n = std::min(cap, e.size());
    std::copy(e.begin(), e.begin() + static_cast<long>(n), out);
    return e.size();
}
int orc_get_rhs(void* s, double* out) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    std::copy(sy.b.begin(), sy.b.end(), out);
    return 0;
}
int orc_get_solution(void* s, double* out) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    std::copy(sy.xsol.begin(), sy.xsol.end(), out);
    return 0;
}
// which: 0 lambda, 1 residual_terms, 2 output(3)
int orc_get_constraint(void* s, size_t id, int which, double* out) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    if (id >= sy.cons.size()) return -1;
    auto& c = sy.cons[id];
    if (which == 0) std::copy(c.lambda.begin(), c.lambda.end(), out);
    else if (which == 1) std::copy(c.res.begin(), c.res.end(), out);
    else if (which == 2) std::copy(c.output.begin(), c.output.end(), out);
    else return -1;
    return 0;
}
// Element evaluation at the current state: residual [N][6], system matrix [N][N][36]
int orc_element_terms(void* s, size_t elem, int update_tangent, double* resid, double* mat) {
    auto* o = static_cast<OrcSystem*>(s);
    if (elem >= o->sys.beams.size()) return -1;
    if (update_tangent) UpdateTangentOperator(o->params, o->sys);
    auto& e = o->sys.beams[elem];
    CalculateQuadraturePointValues(o->params.beta_prime, o->params.gamma_prime, o->sys, e);
    if (resid) std::copy(e.residual.begin(), e.residual.end(), resid);
    if (mat) std::copy(e.sysmat.begin(), e.sysmat.end(), mat);
    return 0;
}
// Element QP setup views: which 0 weight, 1 jacobian, 2 shape_interp, 3 shape_deriv,
// 4 qp_x0, 5 qp_x0_prime, 6 qp_r0, 7 qp_Mstar, 8 qp_Cstar
int orc_element_setup(void* s, size_t elem, int which, double* out) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    if (elem >= sy.beams.size()) return -1;
    auto& e = sy.beams[elem];
    const std::vector<double>* v = nullptr;
    switch (which) {
        case 0: v = &e.weight; break;
        case 1: v = &e.jacobian; break;
        case 2: v = &e.shape_interp; break;
        case 3: v = &e.shape_deriv; break;
        case 4: v = &e.qp_x0; break;
        case 5: v = &e.qp_x0_prime; break;
        case 6: v = &e.qp_r0; break;
        case 7: v = &e.qp_Mstar; break;
        case 8: v = &e.qp_Cstar; break;
        default: return -1;
    }
    std::copy(v->begin(), v->end(), out);
    return 0;
}
void orc_step_parameters(void* s, double* out) {
    auto& p = static_cast<OrcSystem*>(s)->params;
    const double v[] = {p.h, p.alpha_m, p.alpha_f, p.gamma, p.beta, p.gamma_prime, p.beta_prime, p.conditioner};
    std::copy(v, v + 8, out);
}

// --- restart format (src/state/write_state_to_file.hpp:19-58, read_state_from_file.hpp:22-65)
size_t orc_restart_size(void* s) {
    const size_t n = static_cast<OrcSystem*>(s)->sys.num_nodes;
    return sizeof(size_t) + n * (7 * 4 + 6 * 5) * sizeof(double);
}
int orc_write_restart(void* s, unsigned char* out) {
    auto& sy = static_cast<OrcSystem*>(s)->sys;
    size_t n = sy.num_nodes;
    std::memcpy(out, &n, sizeof(size_t));
    size_t off = sizeof(size_t);
    const std::vector<double>* order[] = {&sy.x0, &sy.x, &sy.q_delta, &sy.q_prev, &sy.q,
                                          &sy.v, &sy.vd, &sy.a, &sy.f};
    for (auto* v : order) {
        std::memcpy(out + off, v->data(), v->size() * sizeof(double));
        off += v->size() * sizeof(double);
    }
    return 0;
}

// --- functor entry points for the reference's unit-test known answers
void orc_fn_strain(const double* x0p, const double* up, const double* r, const double* rp, double* out) {
    const auto e = CalculateStrain({x0p[0], x0p[1], x0p[2]}, {up[0], up[1], up[2]},
                                   {r[0], r[1], r[2], r[3]}, {rp[0], rp[1], rp[2], rp[3]});
    std::copy(e.begin(), e.end(), out);
}
void orc_fn_rotate_section(const double* q, const double* C, double* out) {
    Mat6 c{};
    std::copy(C, C + 36, c.begin());
    const auto r = RotateSectionMatrix({q[0], q[1], q[2], q[3]}, c);
    std::copy(r.begin(), r.end(), out);
}
// beams/calculate_{Ouu,Puu,Quu}.hpp
void orc_fn_ouu(const double* C, const double* x0pupSS, const double* Mt, const double* Nt, double* out) {
    put(CalculateOuu(m6(C), m3(x0pupSS), m3(Mt), m3(Nt)), out);
}
void orc_fn_puu(const double* C, const double* x0pupSS, const double* Nt, double* out) {
    put(CalculatePuu(m6(C), m3(x0pupSS), m3(Nt)), out);
}
void orc_fn_quu(const double* C, const double* x0pupSS, const double* Nt, double* out) {
    put(CalculateQuu(m6(C), m3(x0pupSS), m3(Nt)), out);
}
// beams/calculate_force_FC.hpp, calculate_force_FD.hpp, calculate_temporary_variables.hpp
void orc_fn_fc(const double* C, const double* strain, double* out) {
    put(CalculateForceFC(m6(C), {strain[0], strain[1], strain[2], strain[3], strain[4], strain[5]}), out);
}
void orc_fn_fd(const double* x0pupSS, const double* FC, double* out) {
    put(CalculateForceFD(m3(x0pupSS), {FC[0], FC[1], FC[2], FC[3], FC[4], FC[5]}), out);
}
void orc_fn_temporary_variables(const double* x0p, const double* up, double* out) {
    put(CalculateTemporaryVariables(v3(x0p), v3(up)), out);
}
// masses/calculate_mass_matrix_components.hpp, calculate_{inertial,gravity}_force.hpp,
// calculate_gyroscopic_matrix.hpp, calculate_inertia_stiffness_matrix.hpp: the tilde matrices
// are inputs of their own, as in the reference functors
void orc_fn_eta_rho(const double* Muu, double* eta, double* rho) {
    put(CalculateEta(m6(Muu)), eta);
    put(CalculateRho(m6(Muu)), rho);
}
void orc_fn_inertial_force(double m, const double* ud, const double* w, const double* wd, const double* eta,
                           const double* eta_tilde, const double* rho, const double* w_tilde, const double* wd_tilde,
                           double* out) {
    put(CalculateInertialForce(m, v3(ud), v3(w), v3(wd), v3(eta), m3(eta_tilde), m3(rho), m3(w_tilde), m3(wd_tilde)),
        out);
}
void orc_fn_gravity_force(double m, const double* g, const double* eta_tilde, double* out) {
    put(CalculateGravityForce(m, v3(g), m3(eta_tilde)), out);
}
void orc_fn_gyroscopic_matrix(double m, const double* w, const double* eta, const double* rho, const double* w_tilde,
                              double* out) {
    put(CalculateGyroscopicMatrix(m, v3(w), v3(eta), m3(rho), m3(w_tilde)), out);
}
void orc_fn_inertia_stiffness_matrix(double m, const double* ud, const double* w, const double* wd, const double* eta,
                                     const double* rho, const double* w_tilde, const double* wd_tilde, double* out) {
    put(CalculateInertiaStiffnessMatrix(m, v3(ud), v3(w), v3(wd), v3(eta), m3(rho), m3(w_tilde), m3(wd_tilde)), out);
}
// beams/integrate_{residual_vector,stiffness_matrix,inertia_matrix}.hpp: one element, N nodes,
// Q points; shape tables [N][Q], QP matrices [Q][36], QP forces [Q][6]
void orc_fn_integrate_residual(size_t N, size_t Q, const double* w, const double* J, const double* phi,
                               const double* dphi, const double* Fc, const double* Fd, const double* Fi,
                               const double* Fe, const double* Fg, const double* FX, double* out) {
    IntegrateResidualVector(N, Q, w, J, phi, dphi, Fc, Fd, Fi, Fe, Fg, FX, out);
}
void orc_fn_integrate_stiffness(size_t N, size_t Q, const double* w, const double* J, const double* phi,
                                const double* dphi, const double* Kuu, const double* Puu, const double* Cuu,
                                const double* Ouu, const double* Quu, double* out) {
    std::vector<Mat6> K(Q), P(Q), C(Q), O(Q), Qm(Q);
    for (size_t q = 0; q < Q; ++q) {
        K[q] = m6(Kuu + 36 * q);
        P[q] = m6(Puu + 36 * q);
        C[q] = m6(Cuu + 36 * q);
        O[q] = m6(Ouu + 36 * q);
        Qm[q] = m6(Quu + 36 * q);
    }
    IntegrateStiffnessMatrix(N, Q, w, J, phi, dphi, K.data(), P.data(), C.data(), O.data(), Qm.data(), out);
}
void orc_fn_integrate_inertia(size_t N, size_t Q, const double* w, const double* J, const double* phi,
                              double beta_prime, double gamma_prime, const double* Muu, const double* Guu,
                              double* out) {
    std::vector<Mat6> M(Q), G(Q);
    for (size_t q = 0; q < Q; ++q) {
        M[q] = m6(Muu + 36 * q);
        G[q] = m6(Guu + 36 * q);
    }
    IntegrateInertiaMatrix(N, Q, w, J, phi, beta_prime, gamma_prime, M.data(), G.data(), out);
}
// state/*.hpp, one node
void orc_fn_next_state(double h, double alpha_f, double alpha_m, double beta, double gamma, double* qd, double* v,
                       double* vd, double* a) {
    CalculateNextStateNode(h, alpha_f, alpha_m, beta, gamma, qd, v, vd, a);
}
void orc_fn_displacement(double h, const double* qd, const double* q_prev, double* q) {
    CalculateDisplacementNode(h, qd, q_prev, q);
}
void orc_fn_dynamic_prediction(double h, double beta_prime, double gamma_prime, const double* x, double* qd, double* v,
                               double* vd) {
    UpdateDynamicPredictionNode(h, beta_prime, gamma_prime, x, 6, qd, v, vd);
}
void orc_fn_static_prediction(double h, const double* x, double* qd) { UpdateStaticPredictionNode(h, x, 6, qd); }
void orc_fn_algorithmic_acceleration(double alpha_f, double alpha_m, const double* vd, double* a) {
    UpdateAlgorithmicAccelerationNode(alpha_f, alpha_m, vd, a);
}
void orc_fn_jacobian(size_t N, size_t Q, const double* node_x0, const double* dphi, double* J, double* x0p) {
    CalculateJacobian(N, Q, node_x0, dphi, J, x0p);
}
// solver/compute_row_ptrs.hpp, compute_col_inds.hpp, create_full_matrix.hpp: the CRS graph from the
// reference's inputs (node_state_indices [n_elem][max_nodes], freedom tables [n_cons][6], row
// ranges [n_cons][2]); row_ptrs_in (may be null) replaces the computed row pointers for the column
// indices, as the reference's ComputeColInds takes them; sort != 0 sorts each row.  Returns nnz.
size_t orc_fn_crs_graph(size_t num_system_dofs, size_t num_dofs, size_t n_nodes, const size_t* active,
                        const size_t* nfmt, size_t n_elem, const size_t* nodes_per_elem, size_t max_nodes,
                        const size_t* nsi, size_t n_cons, const size_t* base_active, const size_t* target_active,
                        const size_t* base_nft, const size_t* target_nft, const size_t* row_range,
                        const size_t* row_ptrs_in, int sort, size_t* row_ptrs_out, size_t* col_inds_out,
                        size_t col_cap) {
    CrsGraphIn g;
    g.num_system_dofs = num_system_dofs;
    g.num_dofs = num_dofs;
    g.active_dofs.assign(active, active + n_nodes);
    g.node_freedom_map_table.assign(nfmt, nfmt + n_nodes);
    for (size_t e = 0; e < n_elem; ++e) g.element_nodes.emplace_back(nsi + e * max_nodes, nsi + e * max_nodes + nodes_per_elem[e]);
    for (size_t c = 0; c < n_cons; ++c) {
        g.base_active_dofs.push_back(base_active[c]);
        g.target_active_dofs.push_back(target_active[c]);
        std::array<size_t, 6> b{}, t{};
        std::copy(base_nft + 6 * c, base_nft + 6 * c + 6, b.begin());
        std::copy(target_nft + 6 * c, target_nft + 6 * c + 6, t.begin());
        g.base_node_freedom_table.push_back(b);
        g.target_node_freedom_table.push_back(t);
        g.row_range.emplace_back(row_range[2 * c], row_range[2 * c + 1]);
    }
    auto rp = row_ptrs_in ? std::vector<size_t>(row_ptrs_in, row_ptrs_in + num_dofs + 1) : ComputeRowPtrs(g);
    auto ci = ComputeColInds(g, rp);
    if (sort) SortCrsGraph(rp, ci);
    std::copy(rp.begin(), rp.end(), row_ptrs_out);
    std::copy(ci.begin(), ci.begin() + static_cast<long>(std::min(col_cap, ci.size())), col_inds_out);
    return ci.size();
}
// The assembled system matrix (before conditioning and solving: after orc_debug_assemble) in the
// global dof order, dense n x n row-major, from the independent blocks
void orc_get_matrix(void* h, double* out) {
    const auto& s = static_cast<OrcSystem*>(h)->sys;
    const size_t n = s.num_dofs;
    std::fill(out, out + n * n, 0.);
    for (const auto& blk : s.blocks)
        for (size_t i = 0; i < blk.n; ++i)
            for (size_t j = 0; j < blk.n; ++j) out[blk.gdof[i] * n + blk.gdof[j]] = blk.A[i * blk.n + j];
}
// The full matrix's sorted CRS graph of a system (CreateFullMatrix): returns nnz; outputs may be null
size_t orc_crs_graph(void* h, size_t* row_ptrs, size_t* col_inds, size_t cap) {
    const auto& s = static_cast<OrcSystem*>(h)->sys;
    const auto g = CrsGraphInput(s);
    auto rp = ComputeRowPtrs(g);
    auto ci = ComputeColInds(g, rp);
    SortCrsGraph(rp, ci);
    if (row_ptrs) std::copy(rp.begin(), rp.end(), row_ptrs);
    if (col_inds) std::copy(ci.begin(), ci.begin() + static_cast<long>(std::min(cap, ci.size())), col_inds);
    return ci.size();
}
void orc_fn_global_position(const double* q, const double* x0, double* x) { UpdateGlobalPositionNode(q, x0, x); }
void orc_fn_tangent_operator(double h, const double* qd, double* out) {
    const auto t = CalculateTangentOperatorT(h, qd);
    std::copy(t.begin(), t.end(), out);
}
void orc_fn_rv_to_quat(const double* rv, double* out) {
    const auto q = RotationVectorToQuaternion({rv[0], rv[1], rv[2]});
    std::copy(q.begin(), q.end(), out);
}
void orc_fn_quat_to_rv(const double* q, double* out) {
    const auto r = QuaternionToRotationVector({q[0], q[1], q[2], q[3]});
    std::copy(r.begin(), r.end(), out);
}
void orc_fn_quat_compose(const double* a, const double* b, double* out) {
    const auto q = QuaternionCompose({a[0], a[1], a[2], a[3]}, {b[0], b[1], b[2], b[3]});
    std::copy(q.begin(), q.end(), out);
}
void orc_fn_gauss_legendre(size_t n, double* xi, double* w) {
    std::vector<double> x, ww;
    GaussLegendreRule(n, x, ww);
    std::copy(x.begin(), x.end(), xi);
    std::copy(ww.begin(), ww.end(), w);
}
void orc_fn_gll_points(size_t order, double* out) {
    const auto p = GenerateGLLPoints(order);
    std::copy(p.begin(), p.end(), out);
}
void orc_fn_lagrange(double x, size_t n, const double* xs, double* w, double* dw) {
    std::vector<double> v(xs, xs + n), a, b;
    LagrangePolynomialInterpWeights(x, v, a);
    LagrangePolynomialDerivWeights(x, v, b);
    std::copy(a.begin(), a.end(), w);
    std::copy(b.begin(), b.end(), dw);
}
int orc_fn_lu_solve(size_t n, double* A, double* b) {
    std::vector<int> piv(n);
    const bool ok = LUFactor(n, A, piv.data());
    LUSolve(n, A, piv.data(), b);
    return ok ? 0 : 1;
}

}  // extern "C"
