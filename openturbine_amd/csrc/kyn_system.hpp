// Device-resident system layout shared by the host setup (kyn_host.cpp) and the kernels
// (kyn_kernels.hip).  All arrays are structure-of-arrays in HBM, grouped so that one
// workgroup reads one element's / one block's data contiguously.
//
// Vocabulary (reference names):
//   model      one Model of the reference: its own Step, Newton loop and convergence_err
//   block      one independent connected component of a model's KKT system (blade + root BC)
//   element    one beam element (src/elements/beams/beams.hpp)
//   node       one State row (src/state/state.hpp)
//   constraint one Constraints row (src/constraints/constraints.hpp)
#pragma once

#include <cstdint>

namespace kyn {

constexpr int kMaxElemNodes = 21;  // register tile of the element contraction (3 x 21)
constexpr int kMaxElemQps = 64;
constexpr int kThreads = 256;      // workgroup size of every per-element / per-block kernel
constexpr int kMaxBlockDofs = 143; // register LU tile holds n + 1 <= 144 columns ([A | b])
constexpr int kCondMaxLam = 16;    // condensed solve: constraint rows touching one element
constexpr int kTreeMaxBlock = 6;   // tree interface solve: rows of one vertex (constraint / free node)
constexpr int kTreeMaxVerts = 128; // tree interface solve: vertices of one block's interface (LDS)
// scratch of the no-allocation element-terms entry point: residual [N][6] + matrix [N][N][36]
constexpr int kElemTermsScratch = kMaxElemNodes * 6 + kMaxElemNodes * kMaxElemNodes * 36;
constexpr int kMaxScratchDevices = 64;

// Constraint types (src/constraints/constraint_type.hpp:9-21)
constexpr int kConFixedBC = 1, kConPrescribedBC = 2, kConRigidJoint = 3, kConRevoluteJoint = 4,
              kConRotationControl = 5, kConFixedBC3DOFs = 6, kConPrescribedBC3DOFs = 7,
              kConRigidJoint6DOFsTo3DOFs = 8;
// constraints between a base and a target node (constraint_type.hpp:28-33)
constexpr bool con_two_node(int type) {
    return type == kConRigidJoint || type == kConRevoluteJoint || type == kConRotationControl ||
           type == kConRigidJoint6DOFsTo3DOFs;
}
// constraints on a target with three translational dofs (constraint_type.hpp:64-70)
constexpr bool con_three_dof(int type) {
    return type == kConFixedBC3DOFs || type == kConPrescribedBC3DOFs || type == kConRigidJoint6DOFsTo3DOFs;
}
constexpr bool con_prescribed(int type) { return type == kConPrescribedBC || type == kConPrescribedBC3DOFs; }
// constraints whose input(0) is the control signal at UpdateViews (constraints.hpp:346-353)
constexpr bool con_controlled(int type) { return type == kConRevoluteJoint || type == kConRotationControl; }
// constraint rows (constraint_type.hpp:60-79): the revolute joint frees one rotation, the
// three-dof variants hold the translations only
constexpr int con_num_rows(int type) { return con_three_dof(type) ? 3 : type == kConRevoluteJoint ? 5 : 6; }

struct StepConsts {
    double h, alpha_m, alpha_f, gamma, beta, gamma_prime, beta_prime, conditioner, atol, rtol;
    int is_dynamic;
    int max_iter;
};

struct DevSys {
    StepConsts p;
    int num_models, num_blocks, num_elems, num_nodes, num_cons;
    int precision;      // 0 FP64 tangent, 1 FP32 tangent
    int refine_sweeps;  // FP64 iterative refinement sweeps (FP32 tangent)
    double refine_err_bound;  // refine only in Newton iterations whose predecessor's convergence
                              // error is below this (<= 0: in every iteration)
    int max_block_cons; // most constraints in one block (fused LDS layout)

    // ---- state (src/state/state.hpp), [n][7] / [n][6]
    double *x0, *x, *q_prev, *q, *q_delta, *v, *vd, *a, *f;
    double* tan_rot;  // [n][9]: rotational block of the stored tangent (T^T)

    // ---- models
    const int *model_node_begin, *model_node_end;
    const int *model_block_begin, *model_block_end;
    const int *model_con_begin, *model_con_end;
    const int* model_ndofs;
    const double* model_gravity;  // [m][3]
    const int* model_load_begin;  // into sched_load_nodes (sorted by node), [m+1]
    const int* model_order;       // fused launch: workgroup g runs model model_order[g] (costliest first)

    // ---- blocks
    const int* block_model;
    const int* block_n;           // KKT size
    const int* block_nsys;        // system dofs of the block (6 x nodes)
    const int *block_elem_begin, *block_elem_end;  // into block_elems
    const int* block_elems;
    const int *block_con_begin, *block_con_end;    // into block_cons
    const int* block_cons;
    const int *block_node_begin, *block_node_end;  // into block_nodes (ascending)
    const int* block_nodes;
    const int64_t* block_A_off;   // into A (elements of the storage type)
    const int64_t* block_C_off;   // into Acopy: the refinement's row-major copy of the unfactored
                                  // block, row stride (n + 3) & ~3 (whole 16-byte rows, 16-byte aligned)
    const int* block_vec_off;     // into rhs / xsol

    // ---- node -> block-local dof
    const int* node_block;
    const int* node_local;        // -1 if the node has no dofs
    // 1 for a node with the three translational dofs only (FreedomSignature::JustPosition):
    // its rotational rows in the block are held at x = 0 by a unit diagonal and a zero rhs
    const int* node_pin;
    const int* block_npin;        // pinned nodes in the block (0: nothing to do)
    // 1: the FP32-tangent factorization of the block takes no row interchanges (dynamic solve,
    // every node on a beam element, no translations-only nodes, no masses or springs)
    const int* block_nopiv;

    // ---- elements (src/elements/beams/beams.hpp)
    const int *elem_N, *elem_Q;
    const int* elem_node_off;     // into elem_nodes and node_FX rows
    const int* elem_nodes;
    const int* elem_qp_off;       // into qp arrays (rows)
    const int* elem_shape_off;    // into shape tables (N*Q entries)
    const int64_t* elem_mat_off;  // into elem_mat (N*N*36 entries)
    const double *qp_weight, *qp_jac, *qp_Mstar, *qp_Cstar, *qp_x0p, *qp_r0;
    const double *shape_interp, *shape_deriv;  // [N][Q] per table
    double* node_FX;                           // [sum N][6]
    double* elem_res;                          // [sum N][6]
    double* elem_mat;                          // [sum N^2][36] (plain form)

    // ---- point masses (src/elements/masses/masses.hpp) and springs (src/elements/springs/springs.hpp)
    int num_masses, num_springs;
    const int *block_mass_begin, *block_mass_end, *block_masses;      // per block, in model order
    const int *block_spring_begin, *block_spring_end, *block_springs;
    const int* mass_node;
    const double *mass_Mstar, *mass_r0;  // [k][36] section mass matrix, [k][4] node x0 rotation
    double *mass_res, *mass_mat;         // [k][6], [k][36] (beta' M + gamma' G + K T)
    const int *spring_n1, *spring_n2;
    const double *spring_x0, *spring_lref, *spring_k;  // [k][3] (x0 of node 2 - x0 of node 1), [k], [k]
    double *spring_res, *spring_mat;                   // [k][6] ([f, -f]), [k][9] (a; blocks are +-a)

    // ---- constraints (src/constraints/constraints.hpp)
    const int *con_type, *con_base, *con_target, *con_rows, *con_model_row;
    const int* con_local;  // block-local index of the first constraint row
    const double *con_X0, *con_axes;  // [c][3], [c][9]
    double *con_input, *con_control;  // [c][7], [c]
    double *con_lambda, *con_res, *con_sys_res, *con_output;  // [c][6],[c][6],[c][6],[c][3]
    double *con_tgrad, *con_bgrad, *con_tgradT, *con_bgradT;  // [c][36] (plain form)
    double *con_tlres, *con_blres;                            // [c][6]

    // ---- large blocks (more dofs than one workgroup's LU holds; plain form only): the KKT
    //      matrix is assembled into a CRS graph (the reference's create_full_matrix layout:
    //      per row the node's own dofs, the dofs of every element sharing the node and the
    //      constraint rows on it; constraint rows hold their target and base dofs; columns
    //      ascending) by one workgroup per node / constraint (its columns, contributions in the
    //      oracle's order).  Blocks that do not condense are expanded to a dense column-major
    //      matrix for a library LU.
    unsigned contract_mfma_mask;      // FP32 tangent: bit N set -> element size N contracts on the
                                      // matrix cores, else on the VALU register tile
    const int* block_large;           // 1 if the block takes the large path
    const int64_t* crs_rp_off;        // per block: into crs_rp (n + 1 block-local row pointers)
    const int64_t* crs_nz_off;        // per block: into crs_col / crs_val
    const int* crs_rp;
    const int* crs_col;               // block-local column of each nonzero
    double* crs_val;
    int64_t crs_nnz;                  // nonzeros of all large blocks
    const int64_t* block_big_off;     // into Abig / Abig32 (blocks that do not condense)
    double* Abig;
    float* Abig32;                    // FP32 copy for the mixed-precision factorization
    int num_big_items;
    const int* big_item;              // >= 0: node id, < 0: constraint ~id
    const int *node_elem_begin, *node_elem;      // [n+1], pairs {element, local node index}
    const int *node_mass_begin, *node_mass;      // [n+1], mass ids
    const int *node_spring_begin, *node_spring;  // [n+1], spring ids
    const int *node_con_begin, *node_con;        // [n+1], constraint ids (target or base)

    // ---- condensed solve of large blocks whose elements have disjoint nodes: element dofs are
    //      eliminated per element (LDS LU, one workgroup each), the interface (free nodes and
    //      constraint rows, "z") is solved densely, then element dofs are recovered
    int num_cond_elems;
    const int* cond_elem;             // element id (element order within each block)
    const int* cond_elem_block;
    const int* cond_rows_off;         // [ce+1] into cond_rows (block-local rows of the element's dofs)
    const int* cond_rows;
    const int* cond_lam_off;          // [ce+1] into cond_lam / cond_lam_z
    const int* cond_lam;              // block-local constraint rows touching the element
    const int* cond_lam_z;            // their interface indices
    const int64_t* cond_Y_off;        // into Ybuf: n_e x (m_e + 1), column-major (K_e^-1 [B_e^T | b_e])
    const int64_t* cond_S_off;        // into Sbuf: m_e x (m_e + 1), column-major (B_e K_e^-1 [B_e^T | b_e])
    double *Ybuf, *Sbuf;
    const int* block_z_off;           // [b+1] into cond_z (interface rows of the block, ascending)
    const int* cond_z;
    const int* block_ce_begin;        // [b+1] condensed elements of the block
    const int64_t* block_Z_off;       // into Zbuf (nz x nz column-major)
    double* Zbuf;
    double* zrhs;                     // [sum nz]
    // per block-vector entry of a condensed block: the condensed element owning the dof (-1:
    // interface), the dof's index among that element's rows, and its interface index (-1:
    // eliminated)
    const int* dof_cond_elem;
    const int* dof_cond_local;
    const int* dof_zidx;
    // gather maps of the condensed elements (static: the CRS graph is fixed at setup).  Per
    // element [ce+1] into cond_gat_*: every CRS entry of the element's rows as (block-relative
    // nonzero index, LDS index i * lda + column of [K_e | B_e^T]); per constraint row of an
    // element (cond_lam entry) [l+1] into cond_bt_*: that row's entries in the element's columns
    // as (nonzero index, element-local column)
    const int* cond_gat_off;
    const int* cond_gat_src;
    const int* cond_gat_dst;
    const int* cond_bt_off;
    const int* cond_bt_src;
    const int* cond_bt_col;
    // tree-structured interface systems (block elimination, no fill): per block [6] = first
    // vertex, vertices (0: dense solve), first elimination level, elimination levels, first
    // substitution level, substitution levels; per vertex its rows (interface indices), parent
    // (-1: root) and children; vertex ids by level with level offsets
    const int* block_color;           // per block [2]: first colour (into cond_color_off), colours
    const int *cond_color_off, *cond_color_ce;  // condensed elements grouped by colour (no shared constraint row)
    const int* tree_block;
    const int *tree_vsize, *tree_vrow_off, *tree_vrows, *tree_vparent, *tree_vchild_off, *tree_vchildren;
    const int *tree_elim, *tree_elim_lvl, *tree_back, *tree_back_lvl;

    // ---- solver
    void* A;              // blocks' dense KKT matrices (double, or float for FP32 tangent)
    void* Acopy;          // unfactored copy for refinement (FP32 tangent with sweeps > 0)
    double *rhs, *xsol;   // [sum n]
    double* rhs_raw;      // assembled b before conditioning (for kyn_get_rhs)

    // ---- per-model Newton bookkeeping
    double* err_hist;     // [m][max_iter]
    int* model_iters;
    int* model_status;    // 0 converged, 1 not converged, 2 iterating
    int* model_active;
    int* model_iter_total;  // Newton iterations summed over steps
    int* model_fail_total;  // steps that did not converge
    int* any_active;      // plain form: set by k_update when any model keeps iterating

    // ---- device-resident input schedule
    const double* sched_cdisp;   // [steps][num_cons][7] or null
    const int* sched_load_nodes; // [nl]
    const double* sched_loads;   // [steps][nl][6] or null
    int sched_num_loaded;
    int sched_steps;
    int* sched_slot;             // device counter: schedule row of the next step

    // ---- bounded waits: a wave that waits on another (the blocked LU's look-ahead counter)
    //      gives up after lu_wait_limit polls (< 0: at its first wait), sets bit 0 of
    //      dev_error and goes on, so the kernel still ends; the host reports the flag
    int lu_wait_limit;
    int* dev_error;
};

constexpr int kLuWaitLimitDefault = 1 << 20;  // polls of ~64 cycles: ~28 ms at 2.4 GHz
constexpr int kDevErrLuWait = 1;

}  // namespace kyn
