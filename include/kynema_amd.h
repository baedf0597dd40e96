/*
 * kynema_amd.h — C-ABI of the MI355X (gfx950) beam Newton path.
 *
 * The reference exposes this path as compile-time C++ templates, not as an FFI:
 *   Model::AddNode / AddBeamElement / Add*Constraint      src/model/model.hpp:112-169, 477-541
 *   Model::CreateSystemWithSolver / CreateSolver           src/model/model.hpp:39-63, 647-656
 *   StepParameters(is_dynamic, max_iter, h, rho_inf, ...)  src/step/step_parameters.hpp:37-52
 *   bool Step(parameters, solver, elements, state, constraints)   src/step/step.hpp:42-124
 *   Constraints::UpdateDisplacement / UpdateViews          src/constraints/constraints.hpp:338-353
 *   State views q, v, vd, a, x, q_delta, f                 src/state/state.hpp:17-68
 *   Solver::convergence_err                                src/solver/solver.hpp:83
 *   CopyStateData / CloneState                             src/state/copy_state_data.hpp:20-34
 *   WriteStateToFile / ReadStateFromFile                   src/state/write_state_to_file.hpp:19-58
 * Each entry point below names the one it stands in for.
 *
 * Conventions: opaque handles own all device memory; host arrays are copied in and out
 * (no pointer is retained); nothing throws across the ABI; failures return a negative
 * code and set a thread-local message readable with kyn_last_error().  One HIP stream per
 * system; a system is not thread-safe, distinct systems may be driven concurrently.
 * All arrays are row-major FP64 unless stated.
 */
#ifndef KYNEMA_AMD_H
#define KYNEMA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define KYN_OK 0
#define KYN_NOT_CONVERGED 1 /* Step returned false (src/step/step.hpp:55-57) */
#define KYN_ERR_INVALID (-1)
#define KYN_ERR_HIP (-2)
#define KYN_ERR_UNSUPPORTED (-3)

/* ---- constraint types (src/constraints/constraint_type.hpp:9-21) ---- */
#define KYN_FIXED_BC 1
#define KYN_PRESCRIBED_BC 2
#define KYN_RIGID_JOINT 3
#define KYN_REVOLUTE_JOINT 4
#define KYN_ROTATION_CONTROL 5 /* base rotation composed with a commanded rotation about `axis` */
#define KYN_FIXED_BC_3DOFS 6   /* translations of a three-dof node held */
#define KYN_PRESCRIBED_BC_3DOFS 7
#define KYN_RIGID_JOINT_6DOFS_TO_3DOFS 8 /* a six-dof base carries a three-dof target */

/* ---- execution form ---- */
#define KYN_EXEC_PLAIN 0 /* one kernel per stage, host reads the error every iteration */
#define KYN_EXEC_FUSED 1 /* one kernel per time step, Newton loop on the device */

/* ---- tangent precision ---- */
#define KYN_TANGENT_FP64 0
/* FP32 tangent (QP matrices in FP32, node-pair contraction on the matrix cores), FP64 residual,
 * rotations and update.  Fused form: FP32 LU of each block with FP64 refinement.  Plain form:
 * small blocks factor in FP64; large blocks factor in FP32 with FP64 refinement. */
#define KYN_TANGENT_FP32 1

/* ---- state fields for kyn_get_state / kyn_set_state (src/state/state.hpp:29-38) ---- */
#define KYN_X0 0      /* [n][7] */
#define KYN_X 1       /* [n][7] */
#define KYN_Q_PREV 2  /* [n][7] */
#define KYN_Q 3       /* [n][7] */
#define KYN_Q_DELTA 4 /* [n][6] */
#define KYN_V 5       /* [n][6] */
#define KYN_VD 6      /* [n][6] */
#define KYN_A 7       /* [n][6] */
#define KYN_F 8       /* [n][6] */
#define KYN_TANGENT 9 /* [n][36] */

typedef struct kyn_model_s* kyn_model;
typedef struct kyn_system_s* kyn_system;

typedef struct {
    int is_dynamic_solve; /* StepParameters::is_dynamic_solve */
    size_t max_iter;      /* StepParameters::max_iter */
    double h;             /* time step */
    double rho_inf;       /* generalized-alpha spectral radius */
    double atol;          /* absolute_convergence_tol (reference default 1e-5) */
    double rtol;          /* relative_convergence_tol (reference default 1e-3) */
} kyn_step_params;

typedef struct {
    int exec_mode;     /* KYN_EXEC_PLAIN or KYN_EXEC_FUSED */
    int precision;     /* KYN_TANGENT_FP64 or KYN_TANGENT_FP32 */
    int refine_sweeps; /* FP64 iterative-refinement sweeps of the FP32 solve (0..8) */
    double refine_err_bound; /* fused form: refine only in Newton iterations whose previous
                                convergence error is below this bound (<= 0: every iteration) */
} kyn_exec_options;

/* Last error message of the calling thread (empty if none). */
const char* kyn_last_error(void);
/* Device name and architecture of the current HIP device, e.g. "gfx950". */
int kyn_device_info(char* name, size_t name_cap, char* arch, size_t arch_cap);

/* ---- model building: src/model/model.hpp ---- */
kyn_model kyn_model_create(void);                                 /* Model() */
void kyn_model_destroy(kyn_model m);
int kyn_model_set_gravity(kyn_model m, const double g[3]);        /* Model::SetGravity :97 */
/* Model::AddNode + NodeBuilder (model.hpp:112-116, node.hpp:26-49): returns the node id */
int64_t kyn_model_add_node(kyn_model m, const double x0[7], const double u[7], const double v[6],
                           const double vd[6], double s);
/* Model::AddBeamElement (model.hpp:161-169): sections {s, M*[36], C*[36]}, quadrature {xi, w} */
int64_t kyn_model_add_beam_element(kyn_model m, size_t num_nodes, const size_t* node_ids,
                                   size_t num_sections, const double* section_s,
                                   const double* section_mass, const double* section_stiffness,
                                   size_t num_qps, const double* qp_xi_w);
/* Model::AddMassElement (model.hpp:305-311): 6x6 mass matrix in the node's reference frame,
 * row-major.  Returns the mass element id.  Plain form only (the fused form rejects models
 * that hold masses or springs with KYN_ERR_UNSUPPORTED). */
int64_t kyn_model_add_mass_element(kyn_model m, size_t node, const double mass_matrix[36]);
/* Model::AddSpringElement (model.hpp:376-388): linear spring between two nodes' translational
 * dofs.  Returns the spring element id.  Plain form only; a node held by springs alone has
 * three dofs (FreedomSignature::JustPosition), as in the reference. */
int64_t kyn_model_add_spring_element(kyn_model m, size_t node1, size_t node2, double stiffness,
                                     double undeformed_length);
/* Model::AddFixedBC / AddPrescribedBC / AddRigidJointConstraint / AddRevoluteJointConstraint /
 * AddRotationControl / AddFixedBC3DOFs / AddPrescribedBC3DOFs / AddRigidJoint6DOFsTo3DOFs
 * (model.hpp:477-608). For one-node constraints `base` is ignored. axis and init may be NULL.
 * A rotation control constraint takes its commanded angle (radians about the unit `axis`) from
 * kyn_set_constraint_control.  A node whose only dofs come from springs or from the three-dof
 * constraints has the reference's three translational dofs. */
int64_t kyn_model_add_constraint(kyn_model m, int type, size_t base, size_t target,
                                 const double axis[3], const double init_disp[7]);

/* ---- system: Model::CreateSystemWithSolver (model.hpp:647-656) ----
 * A system advances `num_models` independent models at once (each keeps the reference's own
 * Step semantics: its own Newton loop, convergence test and history).  With num_models == 1
 * it is exactly CreateSystemWithSolver of that model. */
kyn_system kyn_system_create(const kyn_model* models, size_t num_models,
                             const kyn_step_params* params, const kyn_exec_options* options);
void kyn_system_destroy(kyn_system s);
size_t kyn_system_num_models(kyn_system s);
size_t kyn_system_num_nodes(kyn_system s);       /* all models, concatenated in order */
size_t kyn_system_num_constraints(kyn_system s);
size_t kyn_system_num_dofs(kyn_system s, size_t model);   /* Solver::num_dofs */
size_t kyn_system_node_offset(kyn_system s, size_t model);
size_t kyn_system_constraint_offset(kyn_system s, size_t model);

/* ---- stepping: Step (step.hpp:42-124) ----
 * Advances every model by one step.  Returns KYN_OK if every model converged,
 * KYN_NOT_CONVERGED if any did not (per-model status via kyn_get_step_status). */
int kyn_step(kyn_system s);
/* Advances num_steps steps back to back with one synchronisation at the end (the inputs come
 * from kyn_set_input_schedule).  use_graph != 0 replays a captured hipGraph of one fused
 * step.  Returns KYN_OK if every model converged in every step so far.  A loop of Step
 * calls in the reference (test_rotor.cpp:112-135); no single counterpart. */
int kyn_step_many(kyn_system s, size_t num_steps, int use_graph);
/* per model: Newton iterations summed over all steps, and the number of failed steps */
int kyn_get_totals(kyn_system s, int* iterations, int* failures);
/* per model: 0 converged / 1 not converged, and Newton iteration count of the last step */
int kyn_get_step_status(kyn_system s, int* status, int* iterations);
/* Solver::convergence_err of one model for the last step; returns the count */
size_t kyn_get_convergence_history(kyn_system s, size_t model, double* err, size_t cap);
/* Device time (ms) spent in the last kyn_step, measured with HIP events on the system stream */
double kyn_last_step_ms(kyn_system s);
/* Average duration (ms) of the step kernel (fused form) or of the element kernel (plain form)
 * over the last kyn_step, from HIP events recorded around those launches */
double kyn_last_kernel_ms(kyn_system s);

/* ---- per-step inputs ---- */
/* Constraints::UpdateDisplacement (constraints.hpp:338-343); id is system-wide */
int kyn_set_constraint_displacement(kyn_system s, size_t id, const double disp[7]);
int kyn_set_constraint_displacements(kyn_system s, size_t count, const size_t* ids,
                                     const double* disp /* [count][7] */);
/* control signal read at UpdateViews (constraints.hpp:346-353): revolute torque, or the
 * commanded angle of a rotation control constraint */
int kyn_set_constraint_control(kyn_system s, size_t id, double value);
/* State::f nodal loads (blade_interface.cpp:56-61), all nodes [n][6] */
int kyn_set_nodal_loads(kyn_system s, const double* f);
/* Beams::node_FX of one element [N][6] (test_cantilever_beam.cpp:105-107) */
int kyn_set_element_node_loads(kyn_system s, size_t element, const double* fx);
/* Device-resident input schedule: step k (counted from this call) takes
 *   constraint displacements cdisp[k][num_constraints][7] (NULL: unchanged) and
 *   nodal loads loads[k][num_loaded][6] on nodes load_nodes[num_loaded] (NULL: unchanged).
 * Replaces per-step host calls for long runs; no counterpart in the reference. */
int kyn_set_input_schedule(kyn_system s, size_t num_steps, const double* cdisp,
                           size_t num_loaded, const size_t* load_nodes, const double* loads);

/* ---- state access: State views (state.hpp:29-38) ---- */
int kyn_get_state(kyn_system s, int field, double* out);
int kyn_set_state(kyn_system s, int field, const double* in);
/* constraint lambda (6), residual_terms (6) or output (3): which = 0, 1, 2 */
int kyn_get_constraint_values(kyn_system s, size_t id, int which, double* out);
/* Solver::b and Solver::x of one model in the reference's global dof order
 * (b: as assembled, before ConditionR; x: after UnconditionSolution) */
int kyn_get_rhs(kyn_system s, size_t model, double* b);
int kyn_get_solution(kyn_system s, size_t model, double* x);

/* HostState::CopyFromState (src/interfaces/host_state.hpp:48-53) in one transfer: per node
 * x[7], q[7], v[6], vd[6] packed on the device as out[n][26] (the reference makes four
 * device-to-host copies per step). */
int kyn_get_motion(kyn_system s, double* out);

/* ---- state management: CopyStateData (copy_state_data.hpp:20-34), restart files ---- */
int kyn_save_state(kyn_system s);
int kyn_restore_state(kyn_system s);
/* byte image of WriteStateToFile (write_state_to_file.hpp:19-58) for one model */
size_t kyn_restart_size(kyn_system s, size_t model);
int kyn_write_restart(kyn_system s, size_t model, unsigned char* out);
int kyn_read_restart(kyn_system s, size_t model, const unsigned char* in, size_t size);

/* ---- element-level access (tests): first assembly of a step without solving
 * (the sequence of test_solver.cpp:97-112), and the element terms it produced ---- */
int kyn_debug_assemble(kyn_system s);
/* Bound of the one wait between waves (the blocked LU's look-ahead counter): after `polls`
 * polls (default 1 << 20, ~28 ms) the waiting wave gives up, the kernel still ends and the
 * step returns KYN_ERR_HIP; polls < 0 gives up at the first wait (tests of that error path).
 * No counterpart in the reference (its LU runs on the host). */
int kyn_debug_set_wait_limit(kyn_system s, int polls);
/* Engine of the FP32 tangent's node-pair contraction per element node count N: bit N of `mask`
 * set selects the matrix cores, clear the VALU register tile.  The default is the compiled-in
 * tuning table (openturbine_amd/autotune.py); this call overrides it (the autotuner, tests).
 * No counterpart in the reference. */
int kyn_debug_set_contraction_mask(kyn_system s, unsigned mask);
int kyn_get_element_terms(kyn_system s, size_t element, double* residual /* [N][6] */,
                          double* system_matrix /* [N][N][36] */);
/* The CRS matrix of a model whose dofs form one large block (plain form), in the reference's
 * dof order: the graph of Solver's full matrix (src/solver/create_full_matrix.hpp:19-74) and
 * the values of the last assembly (Solver::A after AssembleSystemMatrix, conditioned).  Returns
 * the number of nonzeros (0 on error); writes row_ptrs [num_dofs + 1], col_inds and values
 * [nnz] when cap >= nnz and all three are non-NULL. */
size_t kyn_debug_get_crs(kyn_system s, size_t model, size_t* row_ptrs, size_t* col_inds, double* values, size_t cap);
/* No counterpart in the reference: for callers that cannot allocate device memory (HIP graph
 * capture, serving loops).  One beam element's FP64 residual [N][6] and iteration matrix
 * [N][N][36] at the system's current state -- the values kyn_debug_assemble followed by
 * kyn_get_element_terms gives at that state -- formed in a file-scope device array of the
 * library (one per GPU, 128 016 bytes: N <= 21, larger elements return KYN_ERR_INVALID) and
 * copied to the caller's buffers (host or device memory, NULL to skip).  Allocates nothing.
 * The system's device must be current; calls that overlap on one GPU race for the array. */
int kyn_element_terms_noalloc(kyn_system s, size_t element, double* residual, double* system_matrix);

/* ---- model inspection (host only) ---- */
size_t kyn_model_num_nodes(kyn_model m);
int kyn_model_get_node(kyn_model m, size_t id, double x0[7], double u[7], double v[6], double vd[6], double* s);
/* sizes of beam element `elem`, then its node ids [N], sections {s [S], M* [S][36], C* [S][36]}
 * and quadrature {xi, w} [Q][2]; NULL outputs are skipped */
int kyn_model_beam_element_info(kyn_model m, size_t elem, size_t* num_nodes, size_t* num_sections,
                                size_t* num_qps);
int kyn_model_get_beam_element(kyn_model m, size_t elem, size_t* node_ids, double* section_s,
                               double* section_mass, double* section_stiffness, double* qp_xi_w);

/* ---- blade interface: interfaces::BladeInterface (src/interfaces/blade/blade_interface.hpp:17-66)
 *      built from components::SolutionInput and components::BeamInput
 *      (src/interfaces/components/solution_input.hpp, beam_input.hpp) ---- */
typedef struct kyn_blade_interface_s* kyn_blade_interface;

#define KYN_REF_AXIS_X 0 /* components::ReferenceAxisOrientation::X */
#define KYN_REF_AXIS_Z 1 /* ::Z: reference axis points and section matrices rotated from z onto x */

typedef struct {
    double gravity[3];
    int dynamic_solve;
    double time_step;
    double rho_inf;
    size_t max_iter;
    double absolute_error_tolerance;
    double relative_error_tolerance;
} kyn_solution_input;

typedef struct {
    size_t element_order;              /* nodes - 1 */
    size_t section_refinement;         /* interpolated sections between given ones */
    int reference_axis_orientation;    /* KYN_REF_AXIS_X or KYN_REF_AXIS_Z */
    size_t num_ref_axis_points;
    const double* ref_axis_grid;       /* [k] positions in [0, 1] */
    const double* ref_axis_coordinates;/* [k][3] */
    size_t num_twist_points;
    const double* twist_grid;          /* [t] */
    const double* twist;               /* [t], as BeamInput::ref_axis.twist */
    size_t num_sections;
    const double* section_grid;        /* [s] */
    const double* section_mass;        /* [s][36], row-major */
    const double* section_stiffness;   /* [s][36], row-major */
    int prescribe_root_motion;         /* adds a PrescribedBC on the root node */
    double root_position[7];           /* {x, y, z, qw, qx, qy, qz} */
    double root_velocity[6];
    double root_acceleration[6];
} kyn_beam_input;

/* components::Beam::Beam (beam.cpp:17-24) alone, host only: adds the blade's nodes, its beam
 * element and (if requested) the root PrescribedBC to `m`; returns their ids. */
int kyn_beam_build_model(kyn_model m, const kyn_beam_input* in, int64_t* beam_element_id,
                         int64_t* root_constraint_id);
/* BladeInterface::BladeInterface (blade_interface.cpp:11-45); options may be NULL (fused FP64) */
kyn_blade_interface kyn_blade_interface_create(const kyn_solution_input* solution, const kyn_beam_input* blade,
                                               const kyn_exec_options* options);
void kyn_blade_interface_destroy(kyn_blade_interface bi);
size_t kyn_blade_interface_num_nodes(kyn_blade_interface bi);
/* NodeData::loads of blade node `node` (beam.hpp: nodes[node].loads) */
int kyn_blade_interface_set_node_loads(kyn_blade_interface bi, size_t node, const double loads[6]);
int kyn_blade_interface_get_node_loads(kyn_blade_interface bi, size_t node, double loads[6]);
/* Beam::AddPointLoad (beam.cpp:33-45) and Beam::ClearLoads (:47-51) */
int kyn_blade_interface_add_point_load(kyn_blade_interface bi, double s, const double loads[6]);
int kyn_blade_interface_clear_loads(kyn_blade_interface bi);
/* BladeInterface::Step (blade_interface.cpp:54-82): KYN_OK, or KYN_NOT_CONVERGED (node motion
 * then keeps the previous step's values) */
int kyn_blade_interface_step(kyn_blade_interface bi);
/* NodeData::position / displacement / velocity / acceleration (node_data.cpp:12-21); NULL skips */
int kyn_blade_interface_get_node_motion(kyn_blade_interface bi, size_t node, double position[7],
                                        double displacement[7], double velocity[6], double acceleration[6]);
/* BladeInterface::SaveState / RestoreState (blade_interface.cpp:84-96), SetRootDisplacement (:98-103) */
int kyn_blade_interface_save_state(kyn_blade_interface bi);
int kyn_blade_interface_restore_state(kyn_blade_interface bi);
int kyn_blade_interface_set_root_displacement(kyn_blade_interface bi, const double u[7]);
/* the underlying system (state access, restart files) */
kyn_system kyn_blade_interface_system(kyn_blade_interface bi);

#ifdef __cplusplus
}
#endif

#endif /* KYNEMA_AMD_H */
