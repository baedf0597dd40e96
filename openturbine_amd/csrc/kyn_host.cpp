// Host side of the MI355X beam Newton path: the Model facade, one-time setup on the CPU
// (shape functions, QP interpolation, DOF tables, independent blocks), upload to SoA device
// buffers, the step drivers of both execution forms, and the C-ABI of include/kynema_amd.h.
//
// Setup mirrors src/model/model.hpp (CreateSystem, CreateSolver), src/elements/beams/
// create_beams.hpp + populate_element_views.hpp + calculate_jacobian.hpp, src/constraints/
// constraints.hpp:76-266 and src/dof_management/*.hpp.  It runs once per system.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kynema_amd.h"
#include "kyn_system.hpp"
#include "kyn_tuning.inc"

namespace kyn {
size_t fused_lds_bytes(int Nm, int Qm, int nm, int cm, int refine, int precision);
size_t element_lds_bytes(int Nm, int Qm);
size_t solve_lds_bytes(int nm);
hipError_t set_lds_limits(size_t bytes);
void launch_fused_step(const DevSys& S, int Nm, int Qm, int nm, hipStream_t st);
void launch_advance_slot(const DevSys& S, hipStream_t st);
void launch_plain_predict(const DevSys& S, hipStream_t st);
void launch_plain_iteration(const DevSys& S, int Nm, int Qm, int nm, int iter, bool solve,
                            hipStream_t st, hipEvent_t e0, hipEvent_t e1, bool defer_update);
void launch_plain_update(const DevSys& S, int iter, hipStream_t st);
void launch_large_uncondition(const DevSys& S, int b, int n, hipStream_t st);
void launch_crs_residual(const DevSys& S, int b, int n, const double* x, double* r, hipStream_t st);
void launch_crs_to_dense(const DevSys& S, int b, int n, hipStream_t st);
void launch_tree_interface(const DevSys& S, int b, hipStream_t st);
size_t cond_lds(int ne, int me);
void launch_cond_factor(const DevSys& S, size_t lds, bool fp32, const double* rin, hipStream_t st);
void launch_cond_interface(const DevSys& S, int b, int nz, const int* colour_sizes, int ncolour, const double* rin,
                           hipStream_t st);
void launch_cond_backsub(const DevSys& S, double* xout, hipStream_t st);
void launch_cond_scatter_z(const DevSys& S, int b, int nz, double* xout, hipStream_t st);
void launch_block_axpy_copy(const DevSys& S, int b, int n, const double* d, double* x, double* r, hipStream_t st);
void launch_f64_to_f32(const double* x, float* y, size_t count, hipStream_t st);
void launch_pack_motion(const DevSys& S, double* out, hipStream_t st);
void launch_add_f32_to_f64(const float* d, double* x, int count, hipStream_t st);
void launch_plain_epilogue(const DevSys& S, hipStream_t st);
hipError_t element_terms_scratch(int device, double** ptr);
void launch_element_terms(const DevSys& S, int Nm, int Qm, int e, double* res, double* mat, hipStream_t st);
}  // namespace kyn

namespace {

thread_local std::string g_err;

}  // namespace

namespace kyn {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace kyn

namespace {

using kyn::con_num_rows;
using kyn::con_prescribed;
using kyn::con_three_dof;
using kyn::con_two_node;
using kyn::kConFixedBC;
using kyn::kConRigidJoint6DOFsTo3DOFs;
using kyn::kConRotationControl;

struct KynError : std::runtime_error {
    int code;
    KynError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw KynError(KYN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------------------
// Model (src/model/model.hpp, node.hpp, elements/beams/beam_element.hpp, constraint.hpp)
// ---------------------------------------------------------------------------------------
struct HNode {
    std::array<double, 7> x0, u;
    std::array<double, 6> v, vd;
    double s;
};
struct HSection {
    double s;
    std::array<double, 36> M, C;
};
struct HBeam {
    std::vector<size_t> nodes;
    std::vector<HSection> sections;
    std::vector<std::array<double, 2>> quad;
};
struct HMass {  // Model::AddMassElement (model.hpp:305)
    size_t node;
    std::array<double, 36> M;
};
struct HSpring {  // Model::AddSpringElement (model.hpp:376)
    size_t n1, n2;
    double k, l0;
};
struct HConstraint {
    int type;
    size_t base, target;
    std::array<double, 3> axis;
    std::array<double, 7> init;
};

}  // namespace

struct kyn_model_s {
    std::vector<HNode> nodes;
    std::vector<HBeam> beams;
    std::vector<HMass> masses;
    std::vector<HSpring> springs;
    std::vector<HConstraint> constraints;
    std::array<double, 3> gravity{0., 0., 0.};
};

namespace {

// ---------------------------------------------------------------------------------------
// Setup math (host, once): Lagrange shape functions and derivatives at the QPs
// (src/math/interpolation.hpp:69-123), linear section interpolation (:21-43)
// ---------------------------------------------------------------------------------------
void lagrange_values(double x, const std::vector<double>& xs, std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 1.);
    std::vector<double> dx(n);
    for (size_t m = 0; m < n; ++m) dx[m] = x - xs[m];
    for (size_t j = 0; j < n; ++j)
        for (size_t m = 0; m < n; ++m)
            if (m != j) w[j] *= dx[m] / (xs[j] - xs[m]);
}
void lagrange_derivs(double x, const std::vector<double>& xs, std::vector<double>& w) {
    const size_t n = xs.size();
    w.assign(n, 0.);
    for (size_t i = 0; i < n; ++i) {
        double sum = 0.;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double prod = 1.;
            for (size_t k = 0; k < n; ++k)
                if (k != i && k != j) prod *= (x - xs[k]) / (xs[i] - xs[k]);
            sum += prod / (xs[i] - xs[j]);
        }
        w[i] = sum;
    }
}
void linear_weights(double x, const std::vector<double>& xs, std::vector<double>& w) {
    w.assign(xs.size(), 0.);
    const auto it = std::lower_bound(xs.begin(), xs.end(), x);
    if (it == xs.begin()) { w.front() = 1.; return; }
    if (it == xs.end()) { w.back() = 1.; return; }
    const size_t k = static_cast<size_t>(it - xs.begin());
    const double t = (x - xs[k - 1]) / (xs[k] - xs[k - 1]);
    w[k - 1] = 1. - t;
    w[k] = t;
}
std::array<double, 4> qcompose(const std::array<double, 4>& a, const std::array<double, 4>& b) {
    return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
            a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
            a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
            a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}
std::array<double, 3> qrotate(const std::array<double, 4>& q, const std::array<double, 3>& v) {
    return {(q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3]) * v[0] +
                2. * (q[1] * q[2] - q[0] * q[3]) * v[1] + 2. * (q[1] * q[3] + q[0] * q[2]) * v[2],
            2. * (q[1] * q[2] + q[0] * q[3]) * v[0] +
                (q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3]) * v[1] +
                2. * (q[2] * q[3] - q[0] * q[1]) * v[2],
            2. * (q[1] * q[3] - q[0] * q[2]) * v[0] + 2. * (q[2] * q[3] + q[0] * q[1]) * v[1] +
                (q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]) * v[2]};
}

size_t uf_root(std::vector<size_t>& p, size_t i) {
    while (p[i] != i) {
        p[i] = p[p[i]];
        i = p[i];
    }
    return i;
}

// Device allocation helper: every buffer is a separate hipMalloc, freed with the system
struct Arena {
    std::vector<void*> ptrs;
    template <typename T>
    T* up(const std::vector<T>& h) {
        T* d = nullptr;
        const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
        hip_check(hipMalloc(&d, bytes), "hipMalloc");
        ptrs.push_back(d);
        if (!h.empty()) hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
        return d;
    }
    template <typename T>
    T* zeros(size_t n) {
        T* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        ptrs.push_back(d);
        hip_check(hipMemset(d, 0, std::max<size_t>(n, 1) * sizeof(T)), "hipMemset");
        return d;
    }
    ~Arena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// System
// ---------------------------------------------------------------------------------------
struct kyn_system_s {
    kyn::DevSys S{};
    Arena arena;
    hipStream_t stream{};
    hipEvent_t ev0{}, ev1{}, ek0{}, ek1{};
    int exec_mode{KYN_EXEC_FUSED};
    int64_t mat_total{0}, A_total{0};
    int Nm{1}, Qm{1}, nm{1};
    int device{0};  // the HIP device current at creation
    float last_step_ms{0.f}, last_kernel_ms{0.f};
    // host copies of the layout
    size_t num_models{}, num_nodes{}, num_cons{}, num_elems{}, num_blocks{};
    std::vector<int> model_node_begin, model_node_end, model_con_begin, model_con_end;
    std::vector<int> model_ndofs, model_nsys;
    std::vector<int> elem_N, elem_node_off;
    std::vector<int> block_n, block_vec_off, block_model;
    std::vector<int> node_block, node_local;
    std::vector<int> con_local, con_rows, con_model_row, con_block;
    std::vector<std::vector<int>> model_gdof_to_block_vec;  // per model: global dof -> rhs index
    // saved state (CopyStateData)
    std::vector<double> saved;
    bool has_saved{false};
    // large blocks (library LU): block ids, sizes, matrix offsets, pivots, scratch vectors
    std::vector<int> large_blocks, large_n, large_nsys;
    std::vector<int64_t> large_off;
    std::vector<int> large_cond, large_nz;  // condensed solve (else dense), interface size
    std::vector<int64_t> crs_rp_off, crs_nz_off;  // per block (large blocks only)
    std::vector<int> tree_nvert;                  // per block: vertices of its interface tree (0: dense solve)
    std::vector<std::vector<int>> colour_sizes;   // per block: condensed elements of each colour
    std::vector<int64_t> large_Z_off;
    size_t cond_lds_bytes{0};
    rocblas_handle blas{};
    rocblas_int *ipiv{}, *info{};
    double *rvec{}, *bvec{};
    double *rfull{}, *dfull{};  // refinement of condensed blocks: residual and correction (rhs layout)
    double* motion{};           // packed x, q, v, vd for kyn_get_motion (allocated on first use)
    float* r32{};
    std::vector<int> model_active_host;
    // graph of one fused step
    hipGraph_t graph{};
    hipGraphExec_t graph_exec{};
    bool graph_ready{false};

    ~kyn_system_s() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ek0) (void)hipEventDestroy(ek0);
        if (ek1) (void)hipEventDestroy(ek1);
        if (blas) (void)rocblas_destroy_handle(blas);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

struct Builder {
    // concatenated host arrays (become DevSys arrays)
    std::vector<double> x0, q, v, vd, a, f;
    std::vector<int> model_node_begin, model_node_end, model_block_begin, model_block_end,
        model_con_begin, model_con_end, model_ndofs;
    std::vector<double> model_gravity;
    std::vector<int> block_model, block_n, block_nsys, block_elem_begin, block_elem_end, block_elems,
        block_con_begin, block_con_end, block_cons, block_node_begin, block_node_end, block_nodes,
        block_vec_off;
    std::vector<int64_t> block_A_off;
    std::vector<int64_t> block_C_off;  // refinement copy (fused form), padded row stride
    int64_t C_total = 0;
    std::vector<int> node_block, node_local, node_pin, block_npin;
    std::vector<int> block_nopiv;  // FP32 tangent: the block factors without row interchanges
    std::vector<int> elem_N, elem_Q, elem_node_off, elem_nodes, elem_qp_off, elem_shape_off;
    std::vector<int64_t> elem_mat_off;
    std::vector<double> qp_weight, qp_jac, qp_Mstar, qp_Cstar, qp_x0p, qp_r0, shape_interp, shape_deriv;
    std::vector<int> con_type, con_base, con_target, con_rows, con_model_row, con_local, con_block;
    std::vector<double> con_X0, con_axes, con_input;
    std::vector<int> block_mass_begin, block_mass_end, block_masses, block_spring_begin, block_spring_end,
        block_springs, mass_node, spring_n1, spring_n2;
    std::vector<double> mass_Mstar, mass_r0, spring_x0, spring_lref, spring_k;
    std::map<std::string, int> shape_tables;
    std::vector<std::vector<int>> gdof_maps;
    std::vector<int> model_nsys;
    int64_t A_total = 0;
    int vec_total = 0;
    int64_t mat_total = 0;
};

// Interface system of a condensed block as a tree of small blocks ("vertices": the rows of one
// constraint, or the six dofs of one free node).  Two vertices are coupled when their rows can
// share an entry of Z = A[z, z] - sum_e B_e K_e^-1 B_e^T: constraints touching one element
// (through S_e), a constraint and a free node it acts on.  When that graph is a forest, block
// elimination from the leaves has no fill: per vertex v, S_v = Z_vv - sum_children Z_vc X_c,
// X_v = S_v^-1 Z_v,parent, zhat_v = S_v^-1 (r_v - sum_children Z_vc zhat_c), then from the roots
// z_v = zhat_v - X_v z_parent.  Vertices of one height (elimination) or depth (substitution)
// are independent.  Returns false (dense solve) when the graph has a cycle, a vertex has more
// than kTreeMaxBlock rows, there are more than kTreeMaxVerts vertices, or a free node would be a
// leaf (a massless one would leave its block singular).
struct TreeSchedule {
    std::vector<int> size, row_off, rows, parent, child_off, children;  // per vertex (block-local ids)
    std::vector<int> elim, elim_lvl, back, back_lvl;                    // vertex ids by level; level offsets
};

bool build_interface_tree(const Builder& B, int b, const std::map<int, int>& owner,
                          const std::vector<std::vector<int>>& elam, const std::vector<int>& zidx, TreeSchedule& T) {
    std::vector<std::vector<int>> vrows;
    std::vector<int> is_free;
    std::map<int, int> vert_of_row;  // block-local constraint row -> vertex
    std::map<int, int> vert_of_free;  // free node -> vertex
    for (int q = B.block_con_begin[b]; q < B.block_con_end[b]; ++q) {
        const int c = B.block_cons[q];
        std::vector<int> r;
        for (int k = 0; k < B.con_rows[c]; ++k) {
            r.push_back(zidx[B.con_local[c] + k]);
            vert_of_row[B.con_local[c] + k] = static_cast<int>(vrows.size());
        }
        vrows.push_back(r);
        is_free.push_back(0);
    }
    for (int q = B.block_node_begin[b]; q < B.block_node_end[b]; ++q) {
        const int node = B.block_nodes[q];
        if (owner.count(node)) continue;
        std::vector<int> r;
        for (int k = 0; k < 6; ++k) r.push_back(zidx[B.node_local[node] + k]);
        vert_of_free[node] = static_cast<int>(vrows.size());
        vrows.push_back(r);
        is_free.push_back(1);
    }
    const int V = static_cast<int>(vrows.size());
    if (V == 0 || V > kyn::kTreeMaxVerts) return false;
    for (const auto& r : vrows)
        if (static_cast<int>(r.size()) > kyn::kTreeMaxBlock) return false;
    std::set<std::pair<int, int>> edges;
    auto edge = [&](int u, int v) {
        if (u != v) edges.insert({std::min(u, v), std::max(u, v)});
    };
    for (const auto& l : elam) {
        std::set<int> vs;
        for (int row : l) vs.insert(vert_of_row.at(row));
        for (int u : vs)
            for (int v : vs) edge(u, v);
    }
    for (int q = B.block_con_begin[b]; q < B.block_con_end[b]; ++q) {
        const int c = B.block_cons[q];
        const int vc = vert_of_row.at(B.con_local[c]);
        const int nodes[2] = {B.con_target[c], con_two_node(B.con_type[c]) ? B.con_base[c] : -1};
        for (int node : nodes)
            if (node >= 0 && vert_of_free.count(node)) edge(vc, vert_of_free[node]);
    }
    // forest check (union-find), adjacency
    std::vector<int> uf(V);
    std::iota(uf.begin(), uf.end(), 0);
    std::function<int(int)> find = [&](int x) { return uf[x] == x ? x : uf[x] = find(uf[x]); };
    std::vector<std::vector<int>> adj(V);
    for (const auto& e : edges) {
        const int a = find(e.first), c = find(e.second);
        if (a == c) return false;  // a cycle: fill-in, not this solver
        uf[a] = c;
        adj[e.first].push_back(e.second);
        adj[e.second].push_back(e.first);
    }
    // roots: per component the free node of highest degree, else the vertex of highest degree
    std::vector<int> parent(V, -2), depth(V, 0), order;
    for (int s = 0; s < V; ++s) {
        if (parent[s] != -2) continue;
        std::vector<int> comp{s};
        for (size_t k = 0; k < comp.size(); ++k)
            for (int u : adj[comp[k]])
                if (std::find(comp.begin(), comp.end(), u) == comp.end()) comp.push_back(u);
        int root = comp[0];
        for (int u : comp) {
            const auto key = [&](int x) { return std::make_pair(is_free[x], static_cast<int>(adj[x].size())); };
            if (key(u) > key(root)) root = u;
        }
        parent[root] = -1;
        std::vector<int> bfs{root};
        for (size_t k = 0; k < bfs.size(); ++k)
            for (int u : adj[bfs[k]])
                if (parent[u] == -2) {
                    parent[u] = bfs[k];
                    depth[u] = depth[bfs[k]] + 1;
                    bfs.push_back(u);
                }
        order.insert(order.end(), bfs.begin(), bfs.end());
    }
    std::vector<std::vector<int>> ch(V);
    for (int v = 0; v < V; ++v)
        if (parent[v] >= 0) ch[parent[v]].push_back(v);
    for (int v = 0; v < V; ++v)
        if (is_free[v] && ch[v].empty() && parent[v] >= 0) return false;  // free-node leaf
    std::vector<int> height(V, 0);
    for (auto it = order.rbegin(); it != order.rend(); ++it)
        for (int c : ch[*it]) height[*it] = std::max(height[*it], height[c] + 1);
    T = TreeSchedule{};
    for (int v = 0; v < V; ++v) {
        T.size.push_back(static_cast<int>(vrows[v].size()));
        T.row_off.push_back(static_cast<int>(T.rows.size()));
        T.rows.insert(T.rows.end(), vrows[v].begin(), vrows[v].end());
        T.parent.push_back(parent[v]);
        T.child_off.push_back(static_cast<int>(T.children.size()));
        T.children.insert(T.children.end(), ch[v].begin(), ch[v].end());
    }
    T.row_off.push_back(static_cast<int>(T.rows.size()));
    T.child_off.push_back(static_cast<int>(T.children.size()));
    auto levels = [&](const std::vector<int>& key, std::vector<int>& ids, std::vector<int>& off) {
        const int L = *std::max_element(key.begin(), key.end()) + 1;
        off.assign(1, 0);
        for (int l = 0; l < L; ++l) {
            for (int v = 0; v < V; ++v)
                if (key[v] == l) ids.push_back(v);
            off.push_back(static_cast<int>(ids.size()));
        }
    };
    levels(height, T.elim, T.elim_lvl);
    levels(depth, T.back, T.back_lvl);
    return true;
}

// CRS graph of one large block's KKT matrix in block-local dofs (the graph of
// create_full_matrix.hpp:19-74: compute_row_ptrs, compute_col_inds, then sort_crs_graph).  A
// node's rows hold its own six dofs, the dofs of every node it shares a beam element or a spring
// with, and the rows of the constraints on it; a constraint's rows hold its target's and base's
// dofs.  Columns ascending and distinct in every row.
void block_crs_graph(const Builder& B, int b, std::vector<int>& rp, std::vector<int>& cols) {
    const int n = B.block_n[b];
    std::vector<std::vector<int>> row(static_cast<size_t>(n));
    auto couple = [&](int u, int v) {  // the 6 x 6 node block: rows of u, columns of v
        const int lu = B.node_local[u], lv = B.node_local[v];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) row[lu + r].push_back(lv + c);
    };
    for (int k = B.block_node_begin[b]; k < B.block_node_end[b]; ++k) couple(B.block_nodes[k], B.block_nodes[k]);
    for (int k = B.block_elem_begin[b]; k < B.block_elem_end[b]; ++k) {
        const int e = B.block_elems[k], N = B.elem_N[e], off = B.elem_node_off[e];
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) couple(B.elem_nodes[off + i], B.elem_nodes[off + j]);
    }
    for (int k = B.block_spring_begin[b]; k < B.block_spring_end[b]; ++k) {
        const int sp = B.block_springs[k];
        couple(B.spring_n1[sp], B.spring_n2[sp]);
        couple(B.spring_n2[sp], B.spring_n1[sp]);
    }
    for (int k = B.block_con_begin[b]; k < B.block_con_end[b]; ++k) {
        const int c = B.block_cons[k], lr = B.con_local[c];
        const int nodes[2] = {B.con_target[c], con_two_node(B.con_type[c]) ? B.con_base[c] : -1};
        for (int node : nodes) {
            if (node < 0) continue;
            const int l = B.node_local[node];
            for (int r = 0; r < B.con_rows[c]; ++r)
                for (int d = 0; d < 6; ++d) {
                    row[lr + r].push_back(l + d);
                    row[l + d].push_back(lr + r);
                }
        }
    }
    rp.assign(static_cast<size_t>(n) + 1, 0);
    cols.clear();
    for (int i = 0; i < n; ++i) {
        auto& v = row[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        cols.insert(cols.end(), v.begin(), v.end());
        rp[i + 1] = static_cast<int>(cols.size());
    }
}

void add_model(Builder& B, const kyn_model_s& mo, bool dynamic) {
    const int node0 = static_cast<int>(B.node_block.size());
    const int con0 = static_cast<int>(B.con_type.size());
    const int elem0 = static_cast<int>(B.elem_N.size());
    const int blk0 = static_cast<int>(B.block_n.size());
    const int mass0 = static_cast<int>(B.mass_node.size());
    const int spring0 = static_cast<int>(B.spring_n1.size());
    const size_t n = mo.nodes.size();
    const int mid = static_cast<int>(B.model_node_begin.size());

    // ---- state (copy_nodes_to_state.hpp:13-60)
    for (const auto& nd : mo.nodes) {
        for (int c = 0; c < 7; ++c) {
            B.x0.push_back(nd.x0[c]);
            B.q.push_back(nd.u[c]);
        }
        for (int c = 0; c < 6; ++c) {
            B.v.push_back(nd.v[c]);
            B.vd.push_back(nd.vd[c]);
            B.a.push_back(nd.vd[c]);
            B.f.push_back(0.);
        }
    }
    for (size_t i = 0; i < n; ++i) {
        B.node_block.push_back(-1);
        B.node_local.push_back(-1);
        B.node_pin.push_back(0);
    }
    // ---- validate
    for (const auto& bm : mo.beams) {
        if (bm.nodes.size() < 2 || bm.nodes.size() > static_cast<size_t>(kyn::kMaxElemNodes))
            throw KynError(KYN_ERR_UNSUPPORTED, "beam element node count must be in [2, 21]");
        if (bm.quad.empty() || bm.quad.size() > static_cast<size_t>(kyn::kMaxElemQps))
            throw KynError(KYN_ERR_UNSUPPORTED, "beam element quadrature size must be in [1, 64]");
        for (auto nid : bm.nodes)
            if (nid >= n) throw KynError(KYN_ERR_INVALID, "beam element references an unknown node");
        if (bm.sections.empty()) throw KynError(KYN_ERR_INVALID, "beam element without sections");
    }
    for (const auto& c : mo.constraints) {
        if (c.type < kConFixedBC || c.type > kConRigidJoint6DOFsTo3DOFs) throw KynError(KYN_ERR_UNSUPPORTED, "unsupported constraint type");
        if (c.target >= n || (con_two_node(c.type) && c.base >= n))
            throw KynError(KYN_ERR_INVALID, "constraint references an unknown node");
    }

    for (const auto& ms : mo.masses)
        if (ms.node >= n) throw KynError(KYN_ERR_INVALID, "mass element references an unknown node");
    for (const auto& sp : mo.springs) {
        if (sp.n1 >= n || sp.n2 >= n) throw KynError(KYN_ERR_INVALID, "spring element references an unknown node");
        if (sp.n1 == sp.n2) throw KynError(KYN_ERR_INVALID, "spring element joins a node to itself");
    }

    // ---- DOF tables (assemble_node_freedom_allocation_table.hpp, compute_node_freedom_map_table.hpp)
    // freedom signature bits: 1 translations, 2 rotations.  Springs and the three-dof constraint
    // targets give a node its translations only (FreedomSignature::JustPosition); the block keeps
    // six rows per node and holds a translations-only node's rotational rows at zero
    constexpr int kPos = 1, kAll = 3;
    std::vector<int> sig(n, 0);
    for (const auto& bm : mo.beams)
        for (auto nid : bm.nodes) sig[nid] |= kAll;
    for (const auto& ms : mo.masses) sig[ms.node] |= kAll;
    for (const auto& sp : mo.springs) {
        sig[sp.n1] |= kPos;
        sig[sp.n2] |= kPos;
    }
    for (const auto& c : mo.constraints) {
        sig[c.target] |= con_three_dof(c.type) ? kPos : kAll;
        if (con_two_node(c.type)) sig[c.base] |= kAll;
    }
    std::vector<int> has(n, 0);
    for (size_t i = 0; i < n; ++i) has[i] = sig[i] != 0;
    std::vector<int> dof_map(n, 0);
    int nsys_ref = 0;  // the reference's system dofs (three for a translations-only node)
    for (size_t i = 0; i < n; ++i) {
        dof_map[i] = nsys_ref;
        nsys_ref += sig[i] == kAll ? 6 : sig[i] == kPos ? 3 : 0;
        if (sig[i] == kPos) B.node_pin[node0 + i] = 1;
    }
    std::vector<int> row_first(mo.constraints.size());
    int nrows = 0;
    for (size_t k = 0; k < mo.constraints.size(); ++k) {
        row_first[k] = nrows;
        nrows += con_num_rows(mo.constraints[k].type);
    }
    const int ndofs = nsys_ref + nrows;  // the reference's count (convergence-error normalization)

    // ---- blocks (connected components)
    std::vector<size_t> parent(n);
    std::iota(parent.begin(), parent.end(), size_t{0});
    auto unite = [&](size_t a, size_t b) {
        a = uf_root(parent, a);
        b = uf_root(parent, b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    };
    for (const auto& bm : mo.beams)
        for (size_t k = 1; k < bm.nodes.size(); ++k) unite(bm.nodes[0], bm.nodes[k]);
    for (const auto& c : mo.constraints)
        if (con_two_node(c.type)) unite(c.base, c.target);
    for (const auto& sp : mo.springs) unite(sp.n1, sp.n2);
    std::map<size_t, int> root_block;
    std::vector<std::vector<int>> bnodes, bcons, belems, bmasses, bsprings;
    for (size_t i = 0; i < n; ++i) {
        if (!has[i]) continue;
        const size_t r = uf_root(parent, i);
        if (!root_block.count(r)) {
            root_block[r] = static_cast<int>(bnodes.size());
            bnodes.emplace_back();
            bcons.emplace_back();
            belems.emplace_back();
            bmasses.emplace_back();
            bsprings.emplace_back();
        }
        bnodes[root_block[r]].push_back(static_cast<int>(i));
    }
    for (size_t k = 0; k < mo.constraints.size(); ++k)
        bcons[root_block[uf_root(parent, mo.constraints[k].target)]].push_back(static_cast<int>(k));
    for (size_t k = 0; k < mo.beams.size(); ++k)
        belems[root_block[uf_root(parent, mo.beams[k].nodes[0])]].push_back(static_cast<int>(k));
    for (size_t k = 0; k < mo.masses.size(); ++k)
        bmasses[root_block[uf_root(parent, mo.masses[k].node)]].push_back(static_cast<int>(k));
    for (size_t k = 0; k < mo.springs.size(); ++k)
        bsprings[root_block[uf_root(parent, mo.springs[k].n1)]].push_back(static_cast<int>(k));

    std::vector<int> gmap(ndofs, -1);  // model dof -> rhs index
    std::vector<int> con_loc(mo.constraints.size(), 0), con_blk(mo.constraints.size(), 0);
    for (size_t bi = 0; bi < bnodes.size(); ++bi) {
        const int gb = blk0 + static_cast<int>(bi);
        int l = 0;
        B.block_node_begin.push_back(static_cast<int>(B.block_nodes.size()));
        int npin = 0;
        for (int nd : bnodes[bi]) {
            B.block_nodes.push_back(node0 + nd);
            B.node_block[node0 + nd] = gb;
            B.node_local[node0 + nd] = l;
            const int active = B.node_pin[node0 + nd] ? 3 : 6;
            for (int c = 0; c < active; ++c) gmap[dof_map[nd] + c] = B.vec_total + l + c;
            npin += B.node_pin[node0 + nd];
            l += 6;
        }
        B.block_npin.push_back(npin);
        {   // every node on a beam element of the block (mass and stiffness on all six dofs), no
            // translations-only node, no point masses or springs, dynamic solve (the mass term
            // keeps the node rows nonsingular): the FP32 factorization may skip row interchanges
            std::vector<char> on_beam(n, 0);
            for (int ek : belems[bi])
                for (auto nid : mo.beams[static_cast<size_t>(ek)].nodes) on_beam[nid] = 1;
            bool ok = dynamic && npin == 0 && bmasses[bi].empty() && bsprings[bi].empty();
            for (int nd : bnodes[bi]) ok = ok && on_beam[static_cast<size_t>(nd)];
            B.block_nopiv.push_back(ok ? 1 : 0);
        }
        B.block_node_end.push_back(static_cast<int>(B.block_nodes.size()));
        const int bnsys = l;
        B.block_con_begin.push_back(static_cast<int>(B.block_cons.size()));
        for (int ck : bcons[bi]) {
            B.block_cons.push_back(con0 + ck);
            con_loc[ck] = l;
            con_blk[ck] = gb;
            const int rows = con_num_rows(mo.constraints[ck].type);
            for (int r = 0; r < rows; ++r) gmap[nsys_ref + row_first[ck] + r] = B.vec_total + l + r;
            l += rows;
        }
        B.block_con_end.push_back(static_cast<int>(B.block_cons.size()));
        B.block_elem_begin.push_back(static_cast<int>(B.block_elems.size()));
        for (int ek : belems[bi]) B.block_elems.push_back(elem0 + ek);
        B.block_elem_end.push_back(static_cast<int>(B.block_elems.size()));
        B.block_mass_begin.push_back(static_cast<int>(B.block_masses.size()));
        for (int ek : bmasses[bi]) B.block_masses.push_back(mass0 + ek);
        B.block_mass_end.push_back(static_cast<int>(B.block_masses.size()));
        B.block_spring_begin.push_back(static_cast<int>(B.block_springs.size()));
        for (int ek : bsprings[bi]) B.block_springs.push_back(spring0 + ek);
        B.block_spring_end.push_back(static_cast<int>(B.block_springs.size()));
        B.block_model.push_back(mid);
        B.block_n.push_back(l);
        B.block_nsys.push_back(bnsys);
        B.block_vec_off.push_back(B.vec_total);
        B.block_A_off.push_back(B.A_total);
        B.vec_total += l;
        B.A_total += static_cast<int64_t>(l) * (l + 1);
    }
    B.gdof_maps.push_back(gmap);
    B.model_nsys.push_back(nsys_ref);

    // ---- elements (create_beams.hpp:26-165)
    for (const auto& bm : mo.beams) {
        const int N = static_cast<int>(bm.nodes.size()), Q = static_cast<int>(bm.quad.size());
        B.elem_N.push_back(N);
        B.elem_Q.push_back(Q);
        B.elem_node_off.push_back(static_cast<int>(B.elem_nodes.size()));
        for (auto nid : bm.nodes) B.elem_nodes.push_back(node0 + static_cast<int>(nid));
        B.elem_qp_off.push_back(static_cast<int>(B.qp_weight.size()));
        B.elem_mat_off.push_back(B.mat_total);
        B.mat_total += static_cast<int64_t>(N) * N * 36;
        std::vector<double> node_xi(N), phi(N * Q), dphi(N * Q), w;
        for (int k = 0; k < N; ++k) node_xi[k] = 2. * mo.nodes[bm.nodes[k]].s - 1.;
        for (int q = 0; q < Q; ++q) {
            lagrange_values(bm.quad[q][0], node_xi, w);
            for (int k = 0; k < N; ++k) phi[k * Q + q] = w[k];
            lagrange_derivs(bm.quad[q][0], node_xi, w);
            for (int k = 0; k < N; ++k) dphi[k * Q + q] = w[k];
        }
        std::string key(reinterpret_cast<const char*>(node_xi.data()), node_xi.size() * sizeof(double));
        for (const auto& qq : bm.quad) key.append(reinterpret_cast<const char*>(&qq[0]), sizeof(double));
        auto it = B.shape_tables.find(key);
        if (it == B.shape_tables.end()) {
            const int off = static_cast<int>(B.shape_interp.size());
            B.shape_interp.insert(B.shape_interp.end(), phi.begin(), phi.end());
            B.shape_deriv.insert(B.shape_deriv.end(), dphi.begin(), dphi.end());
            it = B.shape_tables.emplace(key, off).first;
        }
        B.elem_shape_off.push_back(it->second);
        // sections (populate_element_views.hpp:89-141)
        std::vector<double> sxi(bm.sections.size());
        for (size_t k = 0; k < sxi.size(); ++k) sxi[k] = 2. * bm.sections[k].s - 1.;
        for (int q = 0; q < Q; ++q) {
            B.qp_weight.push_back(bm.quad[q][1]);
            linear_weights(bm.quad[q][0], sxi, w);
            std::array<double, 36> M{}, C{};
            for (size_t k = 0; k < sxi.size(); ++k)
                for (int c = 0; c < 36; ++c) {
                    M[c] += bm.sections[k].M[c] * w[k];
                    C[c] += bm.sections[k].C[c] * w[k];
                }
            B.qp_Mstar.insert(B.qp_Mstar.end(), M.begin(), M.end());
            B.qp_Cstar.insert(B.qp_Cstar.end(), C.begin(), C.end());
            // interpolate_QP_rotation.hpp / calculate_jacobian.hpp
            double d[3] = {0., 0., 0.}, r[4] = {0., 0., 0., 0.};
            for (int k = 0; k < N; ++k) {
                const auto& x0 = mo.nodes[bm.nodes[k]].x0;
                for (int c = 0; c < 3; ++c) d[c] += x0[c] * dphi[k * Q + q];
                for (int c = 0; c < 4; ++c) r[c] += x0[3 + c] * phi[k * Q + q];
            }
            const double len = std::sqrt(std::pow(r[0], 2) + std::pow(r[1], 2) + std::pow(r[2], 2) + std::pow(r[3], 2));
            if (len == 0.) {
                r[0] = 1.;
                r[1] = r[2] = r[3] = 0.;
            } else
                for (double& c : r) c /= len;
            const double J = std::sqrt(std::pow(d[0], 2.) + std::pow(d[1], 2.) + std::pow(d[2], 2.));
            B.qp_jac.push_back(J);
            for (double c : d) B.qp_x0p.push_back(c / J);
            for (double c : r) B.qp_r0.push_back(c);
        }
    }

    // ---- masses (create_masses.hpp:24-60) and springs (create_springs.hpp:24-58)
    for (const auto& ms : mo.masses) {
        B.mass_node.push_back(node0 + static_cast<int>(ms.node));
        B.mass_Mstar.insert(B.mass_Mstar.end(), ms.M.begin(), ms.M.end());
        for (int c = 0; c < 4; ++c) B.mass_r0.push_back(mo.nodes[ms.node].x0[3 + c]);
    }
    for (const auto& sp : mo.springs) {
        B.spring_n1.push_back(node0 + static_cast<int>(sp.n1));
        B.spring_n2.push_back(node0 + static_cast<int>(sp.n2));
        for (int c = 0; c < 3; ++c) B.spring_x0.push_back(mo.nodes[sp.n2].x0[c] - mo.nodes[sp.n1].x0[c]);
        B.spring_lref.push_back(sp.l0);
        B.spring_k.push_back(sp.k);
    }

    // ---- constraints (constraints.hpp:76-266)
    for (size_t k = 0; k < mo.constraints.size(); ++k) {
        const auto& c = mo.constraints[k];
        const bool two = con_two_node(c.type);
        B.con_type.push_back(c.type);
        B.con_base.push_back(two ? node0 + static_cast<int>(c.base) : node0 + static_cast<int>(c.target));
        B.con_target.push_back(node0 + static_cast<int>(c.target));
        B.con_rows.push_back(con_num_rows(c.type));
        B.con_model_row.push_back(row_first[k]);
        B.con_local.push_back(con_loc[k]);
        B.con_block.push_back(con_blk[k]);
        std::array<double, 3> X0{0., 0., 0.};
        if (!con_prescribed(c.type)) {
            const auto& t = mo.nodes[c.target].x0;
            if (!two) X0 = {t[0] - c.axis[0], t[1] - c.axis[1], t[2] - c.axis[2]};
            else {
                const auto& bb = mo.nodes[c.base].x0;
                X0 = {t[0] - bb[0], t[1] - bb[1], t[2] - bb[2]};
            }
        }
        for (double v : X0) B.con_X0.push_back(v);
        std::array<double, 9> ax{};
        if (c.type == 4) {
            const double na = std::sqrt(c.axis[0] * c.axis[0] + c.axis[1] * c.axis[1] + c.axis[2] * c.axis[2]);
            std::array<double, 3> xh{};
            if (na != 0.) xh = {c.axis[0] / na, c.axis[1] / na, c.axis[2] / na};
            else {
                const double nx = std::sqrt(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
                if (nx == 0.) throw KynError(KYN_ERR_INVALID, "Cannot normalize a zero vector");
                xh = {X0[0] / nx, X0[1] / nx, X0[2] / nx};
            }
            const std::array<double, 3> cp{0. * xh[2] - 0. * xh[1], 0. * xh[0] - 1. * xh[2], 1. * xh[1] - 0. * xh[0]};
            const double dp = xh[0];
            const double kk = 1. / (1. + dp);
            ax = {cp[0] * cp[0] * kk + dp, cp[1] * cp[0] * kk + cp[2], cp[2] * cp[0] * kk - cp[1],
                  cp[0] * cp[1] * kk - cp[2], cp[1] * cp[1] * kk + dp, cp[2] * cp[1] * kk + cp[0],
                  cp[0] * cp[2] * kk + cp[1], cp[1] * cp[2] * kk - cp[0], cp[2] * cp[2] * kk + dp};
        } else if (c.type == kConRotationControl) {  // constraints.hpp:324-331: the unit axis
            const double na = std::sqrt(c.axis[0] * c.axis[0] + c.axis[1] * c.axis[1] + c.axis[2] * c.axis[2]);
            if (na == 0.) throw KynError(KYN_ERR_INVALID, "Cannot normalize a zero vector");
            ax[0] = c.axis[0] / na;
            ax[1] = c.axis[1] / na;
            ax[2] = c.axis[2] / na;
        } else {
            ax[0] = c.axis[0];
            ax[1] = c.axis[1];
            ax[2] = c.axis[2];
        }
        for (double v : ax) B.con_axes.push_back(v);
        std::array<double, 7> in{0., 0., 0., 1., 0., 0., 0.};
        if (con_prescribed(c.type)) in = c.init;
        for (double v : in) B.con_input.push_back(v);
    }

    B.model_node_begin.push_back(node0);
    B.model_node_end.push_back(node0 + static_cast<int>(n));
    B.model_block_begin.push_back(blk0);
    B.model_block_end.push_back(static_cast<int>(B.block_n.size()));
    B.model_con_begin.push_back(con0);
    B.model_con_end.push_back(con0 + static_cast<int>(mo.constraints.size()));
    B.model_ndofs.push_back(ndofs);
    for (double g : mo.gravity) B.model_gravity.push_back(g);
}

template <typename F>
int guarded(F&& f) {
    try {
        g_err.clear();
        return f();
    } catch (const KynError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return KYN_ERR_INVALID;
    }
}

double* field_ptr(kyn_system_s* s, int field, int& width) {
    auto& S = s->S;
    switch (field) {
        case KYN_X0: width = 7; return S.x0;
        case KYN_X: width = 7; return S.x;
        case KYN_Q_PREV: width = 7; return S.q_prev;
        case KYN_Q: width = 7; return S.q;
        case KYN_Q_DELTA: width = 6; return S.q_delta;
        case KYN_V: width = 6; return S.v;
        case KYN_VD: width = 6; return S.vd;
        case KYN_A: width = 6; return S.a;
        case KYN_F: width = 6; return S.f;
        default: return nullptr;
    }
}

int zoff_host(const kyn_system_s* s, int b) {
    int off = 0;
    for (size_t k = 0; k < s->large_blocks.size(); ++k) {
        if (s->large_blocks[k] == b) return off;
        off += s->large_cond[k] ? s->large_nz[k] : 0;
    }
    return off;
}

void blas_check(rocblas_status st, const char* what) {
    if (st != rocblas_status_success)
        throw KynError(KYN_ERR_HIP, std::string(what) + " failed: " + rocblas_status_to_string(st));
}

// SolveSystem for the large blocks (src/step/solve_system.hpp:14-39), from the CRS matrix the
// assembly wrote.  Condensed blocks: element eliminations read their rows of the CRS matrix,
// the interface system is solved densely; FP32 option: element factors in FP32 and FP64
// refinement on r = b - A x computed on the CRS matrix.  Other blocks: the CRS matrix expanded
// to a dense column-major one, LU with partial pivoting (rocSOLVER getrf: largest |a_ik|, first
// index on ties, as the small-block LU) and the triangular solves on the conditioned rhs; FP32
// option: the factorization in FP32 and FP64 iterative refinement against the FP64 matrix,
// `refine_sweeps` correction solves.  Then UnconditionSolution.  Blocks of models that are no
// longer iterating are skipped.
void solve_large_blocks(kyn_system_s* s) {
    auto& S = s->S;
    const hipStream_t st = s->stream;
    // condensed blocks: element eliminations (all elements at once), interface LU per block,
    // element back substitution.  FP32 option: element factors in FP32, FP64 refinement sweeps
    // (each re-runs the condensed solve on the FP64 residual r = b - A x).
    if (S.num_cond_elems > 0) {
        const bool fp32 = S.precision != KYN_TANGENT_FP64;
        auto cond_solve = [&](const double* rin, double* xout) {
            kyn::launch_cond_factor(S, s->cond_lds_bytes, fp32, rin, st);
            hip_check(hipGetLastError(), "k_cond_factor launch");
            for (size_t k = 0; k < s->large_blocks.size(); ++k) {
                const int b = s->large_blocks[k], nz = s->large_nz[k];
                if (!s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
                kyn::launch_cond_interface(S, b, nz, s->colour_sizes[b].data(), static_cast<int>(s->colour_sizes[b].size()),
                                           rin, st);
                if (s->tree_nvert[b] > 0) {  // tree-structured interface: block elimination in one launch
                    kyn::launch_tree_interface(S, b, st);
                    hip_check(hipGetLastError(), "k_tree_interface launch");
                    continue;
                }
                double* Z = S.Zbuf + s->large_Z_off[k];
                double* z = S.zrhs + zoff_host(s, b);
                blas_check(rocsolver_dgetrf(s->blas, nz, nz, Z, nz, s->ipiv, s->info), "rocsolver_dgetrf");
                blas_check(rocsolver_dgetrs(s->blas, rocblas_operation_none, nz, 1, Z, nz, s->ipiv, z, nz),
                           "rocsolver_dgetrs");
            }
            kyn::launch_cond_backsub(S, xout, st);
            for (size_t k = 0; k < s->large_blocks.size(); ++k) {
                const int b = s->large_blocks[k];
                if (!s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
                kyn::launch_cond_scatter_z(S, b, s->large_nz[k], xout, st);
            }
        };
        if (!fp32) {
            cond_solve(S.rhs, S.xsol);
        } else {
            cond_solve(S.rhs, S.xsol);
            for (int sweep = 0; sweep < S.refine_sweeps; ++sweep) {
                for (size_t k = 0; k < s->large_blocks.size(); ++k) {
                    const int b = s->large_blocks[k], n = s->large_n[k];
                    if (!s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
                    // r = b - A x on the CRS matrix (FP64)
                    kyn::launch_crs_residual(S, b, n, S.xsol, s->rfull, st);
                }
                cond_solve(s->rfull, s->dfull);
                for (size_t k = 0; k < s->large_blocks.size(); ++k) {
                    const int b = s->large_blocks[k], n = s->large_n[k];
                    if (!s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
                    kyn::launch_block_axpy_copy(S, b, n, s->dfull, S.xsol, nullptr, st);  // x += d
                }
            }
        }
        for (size_t k = 0; k < s->large_blocks.size(); ++k) {
            const int b = s->large_blocks[k];
            if (!s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
            kyn::launch_large_uncondition(S, b, s->large_n[k], st);
        }
    }
    for (size_t k = 0; k < s->large_blocks.size(); ++k) {
        const int b = s->large_blocks[k], n = s->large_n[k];
        if (s->large_cond[k] || !s->model_active_host[s->block_model[b]]) continue;
        double* A = S.Abig + s->large_off[k];
        double* x = S.xsol + s->block_vec_off[b];
        const double* rhs = S.rhs + s->block_vec_off[b];
        hip_check(hipMemsetAsync(A, 0, static_cast<size_t>(n) * n * sizeof(double), st), "memset");
        kyn::launch_crs_to_dense(S, b, n, st);
        if (S.precision == KYN_TANGENT_FP64) {
            hip_check(hipMemcpyAsync(x, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st), "D2D");
            blas_check(rocsolver_dgetrf(s->blas, n, n, A, n, s->ipiv, s->info), "rocsolver_dgetrf");
            blas_check(rocsolver_dgetrs(s->blas, rocblas_operation_none, n, 1, A, n, s->ipiv, x, n), "rocsolver_dgetrs");
        } else {
            float* A32 = S.Abig32 + s->large_off[k];
            kyn::launch_f64_to_f32(A, A32, static_cast<size_t>(n) * n, st);
            blas_check(rocsolver_sgetrf(s->blas, n, n, A32, n, s->ipiv, s->info), "rocsolver_sgetrf");
            // x = A32^-1 b, then refinement sweeps
            hip_check(hipMemsetAsync(x, 0, n * sizeof(double), st), "memset");
            hip_check(hipMemcpyAsync(s->rvec, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st), "D2D");
            for (int sweep = 0; sweep <= S.refine_sweeps; ++sweep) {
                kyn::launch_f64_to_f32(s->rvec, s->r32, static_cast<size_t>(n), st);
                blas_check(rocsolver_sgetrs(s->blas, rocblas_operation_none, n, 1, A32, n, s->ipiv, s->r32, n),
                           "rocsolver_sgetrs");
                kyn::launch_add_f32_to_f64(s->r32, x, n, st);
                if (sweep == S.refine_sweeps) break;
                // r = b - A x (FP64)
                hip_check(hipMemcpyAsync(s->rvec, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st), "D2D");
                const double alpha = -1., beta = 1.;
                blas_check(rocblas_dgemv(s->blas, rocblas_operation_none, n, n, &alpha, A, n, x, 1, &beta, s->rvec, 1),
                           "rocblas_dgemv");
            }
        }
        kyn::launch_large_uncondition(S, b, n, st);
    }
}

// Device error flags raised by a kernel that gave up a bounded wait: reported (and cleared)
// after the step, as an error of the step.
void check_device_errors(kyn_system_s* s) {
    int flags = 0;
    hip_check(hipMemcpy(&flags, s->S.dev_error, sizeof(int), hipMemcpyDeviceToHost), "D2H");
    if (flags == 0) return;
    const int zero = 0;
    hip_check(hipMemcpy(s->S.dev_error, &zero, sizeof(int), hipMemcpyHostToDevice), "H2D");
    throw KynError(KYN_ERR_HIP, "a wave of the blocked LU gave up waiting on its look-ahead counter "
                                "(bound " + std::to_string(s->S.lu_wait_limit) + " polls); the step's results are invalid");
}

int run_step(kyn_system_s* s, bool solve = true) {
    auto& S = s->S;
    hip_check(hipEventRecord(s->ev0, s->stream), "event");
    float kms = 0.f;
    if (s->exec_mode == KYN_EXEC_FUSED && solve) {
        hip_check(hipEventRecord(s->ek0, s->stream), "event");
        kyn::launch_fused_step(S, s->Nm, s->Qm, s->nm, s->stream);
        hip_check(hipGetLastError(), "k_fused_step launch");
        hip_check(hipEventRecord(s->ek1, s->stream), "event");
        kyn::launch_advance_slot(S, s->stream);
    } else {
        kyn::launch_plain_predict(S, s->stream);
        hip_check(hipGetLastError(), "k_plain_predict launch");
        int count = 0;
        const bool large = !s->large_blocks.empty();
        if (large) s->model_active_host.assign(s->num_models, 1);
        for (int iter = 0; iter < S.p.max_iter; ++iter) {
            kyn::launch_plain_iteration(S, s->Nm, s->Qm, s->nm, iter, solve, s->stream, s->ek0, s->ek1, large);
            hip_check(hipGetLastError(), "plain iteration launch");
            if (!solve) break;
            if (large) {
                solve_large_blocks(s);
                kyn::launch_plain_update(S, iter, s->stream);
                hip_check(hipGetLastError(), "k_plain_update launch");
                hip_check(hipMemcpyAsync(s->model_active_host.data(), S.model_active, s->num_models * sizeof(int),
                                         hipMemcpyDeviceToHost, s->stream),
                          "D2H");
            }
            int any = 0;
            hip_check(hipMemcpyAsync(&any, S.any_active, sizeof(int), hipMemcpyDeviceToHost, s->stream), "D2H");
            hip_check(hipStreamSynchronize(s->stream), "sync");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, s->ek0, s->ek1), "elapsed");
            kms += ms;
            ++count;
            if (!any) break;
        }
        if (solve) {
            kyn::launch_plain_epilogue(S, s->stream);
            kyn::launch_advance_slot(S, s->stream);
        }
        if (count > 0) kms /= static_cast<float>(count);
    }
    hip_check(hipEventRecord(s->ev1, s->stream), "event");
    hip_check(hipStreamSynchronize(s->stream), "sync");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, s->ev0, s->ev1), "elapsed");
    s->last_step_ms = ms;
    if (s->exec_mode == KYN_EXEC_FUSED && solve) {
        hip_check(hipEventElapsedTime(&kms, s->ek0, s->ek1), "elapsed");
    }
    s->last_kernel_ms = kms;
    check_device_errors(s);
    if (!solve) return KYN_OK;
    std::vector<int> st(s->num_models);
    hip_check(hipMemcpy(st.data(), S.model_status, st.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H");
    for (int v : st)
        if (v != 0) return KYN_NOT_CONVERGED;
    return KYN_OK;
}

}  // namespace

// =======================================================================================
// C-ABI
// =======================================================================================
extern "C" {

const char* kyn_last_error(void) { return g_err.c_str(); }

int kyn_device_info(char* name, size_t name_cap, char* arch, size_t arch_cap) {
    return guarded([&] {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        if (name && name_cap) {
            std::strncpy(name, prop.name, name_cap - 1);
            name[name_cap - 1] = 0;
        }
        if (arch && arch_cap) {
            std::strncpy(arch, prop.gcnArchName, arch_cap - 1);
            arch[arch_cap - 1] = 0;
        }
        return KYN_OK;
    });
}

kyn_model kyn_model_create(void) { return new kyn_model_s(); }
void kyn_model_destroy(kyn_model m) { delete m; }
int kyn_model_set_gravity(kyn_model m, const double g[3]) {
    if (!m || !g) return KYN_ERR_INVALID;
    m->gravity = {g[0], g[1], g[2]};
    return KYN_OK;
}
int64_t kyn_model_add_node(kyn_model m, const double x0[7], const double u[7], const double v[6],
                           const double vd[6], double s) {
    if (!m || !x0) return KYN_ERR_INVALID;
    HNode n{};
    std::copy(x0, x0 + 7, n.x0.begin());
    if (u) std::copy(u, u + 7, n.u.begin());
    else n.u = {0., 0., 0., 1., 0., 0., 0.};
    if (v) std::copy(v, v + 6, n.v.begin());
    if (vd) std::copy(vd, vd + 6, n.vd.begin());
    n.s = s;
    m->nodes.push_back(n);
    return static_cast<int64_t>(m->nodes.size() - 1);
}
int64_t kyn_model_add_beam_element(kyn_model m, size_t num_nodes, const size_t* node_ids,
                                   size_t num_sections, const double* section_s,
                                   const double* section_mass, const double* section_stiffness,
                                   size_t num_qps, const double* qp_xi_w) {
    if (!m || !node_ids || !section_s || !section_mass || !section_stiffness || !qp_xi_w)
        return KYN_ERR_INVALID;
    HBeam b;
    b.nodes.assign(node_ids, node_ids + num_nodes);
    for (size_t k = 0; k < num_sections; ++k) {
        HSection sec{};
        sec.s = section_s[k];
        std::copy(section_mass + 36 * k, section_mass + 36 * k + 36, sec.M.begin());
        std::copy(section_stiffness + 36 * k, section_stiffness + 36 * k + 36, sec.C.begin());
        b.sections.push_back(sec);
    }
    for (size_t q = 0; q < num_qps; ++q) b.quad.push_back({qp_xi_w[2 * q], qp_xi_w[2 * q + 1]});
    m->beams.push_back(std::move(b));
    return static_cast<int64_t>(m->beams.size() - 1);
}
int64_t kyn_model_add_mass_element(kyn_model m, size_t node, const double mass_matrix[36]) {
    if (!m || !mass_matrix) return KYN_ERR_INVALID;
    HMass e{};
    e.node = node;
    std::copy(mass_matrix, mass_matrix + 36, e.M.begin());
    m->masses.push_back(e);
    return static_cast<int64_t>(m->masses.size() - 1);
}
int64_t kyn_model_add_spring_element(kyn_model m, size_t node1, size_t node2, double stiffness,
                                     double undeformed_length) {
    if (!m) return KYN_ERR_INVALID;
    m->springs.push_back(HSpring{node1, node2, stiffness, undeformed_length});
    return static_cast<int64_t>(m->springs.size() - 1);
}
int64_t kyn_model_add_constraint(kyn_model m, int type, size_t base, size_t target, const double axis[3],
                                 const double init_disp[7]) {
    if (!m) return KYN_ERR_INVALID;
    HConstraint c{};
    c.type = type;
    c.base = base;
    c.target = target;
    c.axis = {0., 0., 0.};
    if (axis) std::copy(axis, axis + 3, c.axis.begin());
    c.init = {0., 0., 0., 1., 0., 0., 0.};
    if (init_disp) std::copy(init_disp, init_disp + 7, c.init.begin());
    m->constraints.push_back(c);
    return static_cast<int64_t>(m->constraints.size() - 1);
}

kyn_system kyn_system_create(const kyn_model* models, size_t num_models, const kyn_step_params* params,
                             const kyn_exec_options* options) {
    kyn_system result = nullptr;
    const int rc = guarded([&] {
        if (!models || num_models == 0 || !params) throw KynError(KYN_ERR_INVALID, "null argument");
        auto sys = std::make_unique<kyn_system_s>();
        int dev_count = 0;
        hip_check(hipGetDeviceCount(&dev_count), "hipGetDeviceCount");
        if (dev_count < 1) throw KynError(KYN_ERR_HIP, "no HIP device");
        Builder B;
        for (size_t k = 0; k < num_models; ++k) add_model(B, *models[k], params->is_dynamic_solve != 0);
        auto& S = sys->S;
        // ---- step parameters (step_parameters.hpp:37-52)
        const double rho = params->rho_inf, h = params->h;
        const bool dyn = params->is_dynamic_solve != 0;
        S.p.h = h;
        S.p.alpha_m = (2. * rho - 1.) / (rho + 1.);
        S.p.alpha_f = rho / (rho + 1.);
        S.p.gamma = 0.5 + S.p.alpha_f - S.p.alpha_m;
        S.p.beta = 0.25 * (S.p.gamma + 0.5) * (S.p.gamma + 0.5);
        S.p.gamma_prime = dyn ? S.p.gamma / (h * S.p.beta) : 0.;
        S.p.beta_prime = dyn ? (1. - S.p.alpha_m) / (h * h * S.p.beta * (1. - S.p.alpha_f)) : 0.;
        S.p.conditioner = S.p.beta * h * h;
        S.p.atol = params->atol;
        S.p.rtol = params->rtol;
        S.p.is_dynamic = dyn ? 1 : 0;
        S.p.max_iter = static_cast<int>(params->max_iter);
        if (S.p.max_iter < 1) throw KynError(KYN_ERR_INVALID, "max_iter must be >= 1");
        sys->exec_mode = options ? options->exec_mode : KYN_EXEC_FUSED;
        S.precision = options ? options->precision : KYN_TANGENT_FP64;
        S.refine_sweeps = options ? options->refine_sweeps : 0;
        S.refine_err_bound = options ? options->refine_err_bound : 0.;
        if (S.refine_sweeps < 0 || S.refine_sweeps > 8) throw KynError(KYN_ERR_INVALID, "refine_sweeps in [0, 8]");

        // ---- sizes and eligibility
        S.num_models = static_cast<int>(num_models);
        S.num_blocks = static_cast<int>(B.block_n.size());
        S.num_elems = static_cast<int>(B.elem_N.size());
        S.num_nodes = static_cast<int>(B.node_block.size());
        S.num_cons = static_cast<int>(B.con_type.size());
        int Nm = 2, Qm = 1, nm = 1;
        for (int N : B.elem_N) Nm = std::max(Nm, N);
        for (int Q : B.elem_Q) Qm = std::max(Qm, Q);
        // blocks above one workgroup's LU take the large path (plain form): column-major HBM
        // matrix, library LU; the others keep the per-block LDS factorization
        std::vector<int> block_large(B.block_n.size(), 0);
        int64_t A_small = 0;
        B.block_C_off.assign(B.block_n.size(), 0);
        B.C_total = 0;
        for (size_t b = 0; b < B.block_n.size(); ++b) {
            const int n = B.block_n[b];
            if (n > kyn::kMaxBlockDofs) {
                if (sys->exec_mode == KYN_EXEC_FUSED)
                    throw KynError(KYN_ERR_UNSUPPORTED, "the fused form takes independent KKT blocks of at most " +
                                                            std::to_string(kyn::kMaxBlockDofs) +
                                                            " dofs; use the plain form for larger structures");
                block_large[b] = 1;
                sys->large_blocks.push_back(static_cast<int>(b));
                sys->large_n.push_back(n);
                sys->large_nsys.push_back(B.block_nsys[b]);
                sys->large_off.push_back(0);  // dense offset: set once the condensed blocks are known
                B.block_A_off[b] = 0;
            } else {
                nm = std::max(nm, n);
                B.block_A_off[b] = A_small;
                A_small += static_cast<int64_t>(n) * (n + 1);
                B.block_C_off[b] = B.C_total;  // stays a multiple of 4 elements
                B.C_total += static_cast<int64_t>(n) * ((n + 3) & ~3);
            }
        }
        B.A_total = A_small;
        sys->Nm = Nm;
        sys->Qm = Qm;
        sys->nm = nm;
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        sys->device = dev;
        S.contract_mfma_mask = kyn::kContractMfmaMask;
        int max_lds = 0;
        hip_check(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attr");
        size_t need = 0;
        S.num_masses = static_cast<int>(B.mass_node.size());
        S.num_springs = static_cast<int>(B.spring_n1.size());
        if (sys->exec_mode == KYN_EXEC_FUSED) {
            if (S.num_masses + S.num_springs > 0)
                throw KynError(KYN_ERR_UNSUPPORTED, "the fused form takes beam elements only; use the plain form "
                                                    "for point masses and springs");
            for (int b = 0; b < S.num_blocks; ++b) {
                if (B.block_elem_end[b] - B.block_elem_begin[b] != 1)
                    throw KynError(KYN_ERR_UNSUPPORTED, "the fused form needs exactly one beam element per independent block");
                if (B.block_con_end[b] - B.block_con_begin[b] > 8)
                    throw KynError(KYN_ERR_UNSUPPORTED, "the fused form allows at most 8 constraints per block");
                S.max_block_cons = std::max(S.max_block_cons, B.block_con_end[b] - B.block_con_begin[b]);
            }
            need = kyn::fused_lds_bytes(Nm, Qm, nm, S.max_block_cons, S.refine_sweeps, S.precision);
        } else {
            need = std::max(kyn::element_lds_bytes(Nm, Qm), kyn::solve_lds_bytes(nm));
        }
        if (need > static_cast<size_t>(max_lds))
            throw KynError(KYN_ERR_UNSUPPORTED, "problem needs " + std::to_string(need) + " bytes of LDS, device offers " +
                                                    std::to_string(max_lds));
        hip_check(kyn::set_lds_limits(static_cast<size_t>(max_lds)), "hipFuncSetAttribute");

        // ---- upload
        auto& Ar = sys->arena;
        const size_t n = B.node_block.size();
        S.x0 = Ar.up(B.x0);
        S.x = Ar.zeros<double>(n * 7);
        S.q_prev = Ar.up(B.q);
        S.q = Ar.up(B.q);
        S.q_delta = Ar.zeros<double>(n * 6);
        S.v = Ar.up(B.v);
        S.vd = Ar.up(B.vd);
        S.a = Ar.up(B.a);
        S.f = Ar.up(B.f);
        S.tan_rot = Ar.zeros<double>(n * 9);
        S.model_node_begin = Ar.up(B.model_node_begin);
        S.model_node_end = Ar.up(B.model_node_end);
        S.model_block_begin = Ar.up(B.model_block_begin);
        S.model_block_end = Ar.up(B.model_block_end);
        S.model_con_begin = Ar.up(B.model_con_begin);
        S.model_con_end = Ar.up(B.model_con_end);
        S.model_ndofs = Ar.up(B.model_ndofs);
        S.model_gravity = Ar.up(B.model_gravity);
        S.model_load_begin = Ar.zeros<int>(num_models + 1);
        {
            // dispatch order of the fused launch: longest expected step first (sum over the model's
            // blocks of n^2, between the panel's O(n) and the trailing update's O(n^3)), so the
            // short models fill the tail of the grid
            std::vector<double> cost(num_models, 0.);
            for (size_t b = 0; b < B.block_n.size(); ++b)
                cost[B.block_model[b]] += static_cast<double>(B.block_n[b]) * B.block_n[b];
            std::vector<int> order(num_models);
            for (size_t m = 0; m < num_models; ++m) order[m] = static_cast<int>(m);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
            S.model_order = Ar.up(order);
        }
        S.block_model = Ar.up(B.block_model);
        S.block_n = Ar.up(B.block_n);
        S.block_nsys = Ar.up(B.block_nsys);
        S.block_elem_begin = Ar.up(B.block_elem_begin);
        S.block_elem_end = Ar.up(B.block_elem_end);
        S.block_elems = Ar.up(B.block_elems);
        S.block_con_begin = Ar.up(B.block_con_begin);
        S.block_con_end = Ar.up(B.block_con_end);
        S.block_cons = Ar.up(B.block_cons);
        S.block_node_begin = Ar.up(B.block_node_begin);
        S.block_node_end = Ar.up(B.block_node_end);
        S.block_nodes = Ar.up(B.block_nodes);
        S.block_A_off = Ar.up(B.block_A_off);
        S.block_C_off = Ar.up(B.block_C_off);
        S.block_large = Ar.up(block_large);
        {
            // CRS graphs of the large blocks; the dense matrices of blocks that do not condense are
            // allocated below, once the condensed blocks are known
            sys->crs_rp_off.assign(B.block_n.size(), 0);
            sys->crs_nz_off.assign(B.block_n.size(), 0);
            std::vector<int> rp_all, col_all;  // kept for the condensed elements' gather maps
            {
                std::vector<int> rp, cols;
                for (int b : sys->large_blocks) {
                    block_crs_graph(B, b, rp, cols);
                    sys->crs_rp_off[b] = static_cast<int64_t>(rp_all.size());
                    sys->crs_nz_off[b] = static_cast<int64_t>(col_all.size());
                    rp_all.insert(rp_all.end(), rp.begin(), rp.end());
                    col_all.insert(col_all.end(), cols.begin(), cols.end());
                }
                S.crs_rp_off = Ar.up(sys->crs_rp_off);
                S.crs_nz_off = Ar.up(sys->crs_nz_off);
                S.crs_rp = Ar.up(rp_all);
                S.crs_col = Ar.up(col_all);
                S.crs_nnz = static_cast<int64_t>(col_all.size());
                S.crs_val = Ar.zeros<double>(std::max<size_t>(col_all.size(), 1));
            }
            // work items (nodes, then constraints, of each large block) and node incidence lists
            std::vector<int> items;
            for (int b : sys->large_blocks) {
                for (int k = B.block_node_begin[b]; k < B.block_node_end[b]; ++k) items.push_back(B.block_nodes[k]);
                for (int k = B.block_con_begin[b]; k < B.block_con_end[b]; ++k) items.push_back(~B.block_cons[k]);
            }
            S.num_big_items = static_cast<int>(items.size());
            S.big_item = Ar.up(items);
            const size_t nn = B.node_block.size();
            std::vector<std::vector<int>> ne(nn), nmass(nn), nspr(nn), ncon(nn);
            for (size_t e = 0; e < B.elem_N.size(); ++e)
                for (int i = 0; i < B.elem_N[e]; ++i) {
                    auto& v = ne[B.elem_nodes[B.elem_node_off[e] + i]];
                    v.push_back(static_cast<int>(e));
                    v.push_back(i);
                }
            for (size_t k = 0; k < B.mass_node.size(); ++k) nmass[B.mass_node[k]].push_back(static_cast<int>(k));
            for (size_t k = 0; k < B.spring_n1.size(); ++k) {
                nspr[B.spring_n1[k]].push_back(static_cast<int>(k));
                nspr[B.spring_n2[k]].push_back(static_cast<int>(k));
            }
            for (size_t k = 0; k < B.con_type.size(); ++k) {
                ncon[B.con_target[k]].push_back(static_cast<int>(k));
                if (con_two_node(B.con_type[k]) && B.con_base[k] != B.con_target[k])
                    ncon[B.con_base[k]].push_back(static_cast<int>(k));
            }
            auto csr = [&](const std::vector<std::vector<int>>& lists, int width, std::vector<int>& begin,
                           std::vector<int>& flat) {
                begin.assign(nn + 1, 0);
                for (size_t i = 0; i < nn; ++i) {
                    begin[i + 1] = begin[i] + static_cast<int>(lists[i].size()) / width;
                    flat.insert(flat.end(), lists[i].begin(), lists[i].end());
                }
            };
            std::vector<int> b1, f1, b2, f2, b3, f3, b4, f4;
            csr(ne, 2, b1, f1);
            csr(nmass, 1, b2, f2);
            csr(nspr, 1, b3, f3);
            csr(ncon, 1, b4, f4);
            S.node_elem_begin = Ar.up(b1);
            S.node_elem = Ar.up(f1);
            S.node_mass_begin = Ar.up(b2);
            S.node_mass = Ar.up(f2);
            S.node_spring_begin = Ar.up(b3);
            S.node_spring = Ar.up(f3);
            S.node_con_begin = Ar.up(b4);
            S.node_con = Ar.up(f4);
            // condensed solve tables (FP64 factorization, no springs, elements with disjoint nodes)
            {
                std::vector<int> ce_elem, ce_block, rows_off{0}, rows, lam_off{0}, lam, lam_z, z_off(B.block_n.size() + 1, 0),
                    zl, ce_begin(B.block_n.size() + 1, 0);
                std::vector<int64_t> Y_off, S_off, Z_off(B.block_n.size(), 0);
                int64_t Y_tot = 0, S_tot = 0, Z_tot = 0;
                std::vector<int> dof_ce(static_cast<size_t>(B.vec_total), -1), dof_cl(static_cast<size_t>(B.vec_total), -1),
                    dof_z(static_cast<size_t>(B.vec_total), -1);
                std::vector<int> tree_block(6 * B.block_n.size(), 0), tv_size, tv_row_off, tv_rows, tv_parent, tv_child_off,
                    tv_children, t_elim, t_elim_lvl, t_back, t_back_lvl;
                std::vector<int> block_color(2 * B.block_n.size(), 0), color_off{0}, color_ce;
                std::vector<int> nz_of(B.block_n.size(), 0), cond_of(B.block_n.size(), 0), nce_of(B.block_n.size(), 0);
                std::vector<int> gat_off{0}, gat_src, gat_dst, bt_off{0}, bt_src, bt_col;
                for (int b : sys->large_blocks) {
                    bool ok = true;
                    for (size_t k = 0; k < B.spring_n1.size() && ok; ++k)
                        if (B.node_block[B.spring_n1[k]] == b) ok = false;
                    std::map<int, int> owner;  // node -> element
                    for (int k = B.block_elem_begin[b]; k < B.block_elem_end[b] && ok; ++k) {
                        const int e = B.block_elems[k];
                        if (6 * B.elem_N[e] > kyn::kMaxBlockDofs) ok = false;
                        for (int i = 0; i < B.elem_N[e] && ok; ++i)
                            if (!owner.emplace(B.elem_nodes[B.elem_node_off[e] + i], e).second) ok = false;
                    }
                    std::vector<std::vector<int>> elam;
                    size_t need = 0;
                    if (ok) {
                        for (int k = B.block_elem_begin[b]; k < B.block_elem_end[b]; ++k) {
                            const int e = B.block_elems[k];
                            std::vector<int> l;
                            for (int q = B.block_con_begin[b]; q < B.block_con_end[b]; ++q) {
                                const int c = B.block_cons[q];
                                const bool two = con_two_node(B.con_type[c]);
                                auto in_e = [&](int node) {
                                    auto it = owner.find(node);
                                    return it != owner.end() && it->second == e;
                                };
                                if (in_e(B.con_target[c]) || (two && in_e(B.con_base[c])))
                                    for (int r = 0; r < B.con_rows[c]; ++r) l.push_back(B.con_local[c] + r);
                            }
                            if (static_cast<int>(l.size()) > kyn::kCondMaxLam) ok = false;
                            need = std::max(need, kyn::cond_lds(6 * B.elem_N[e], static_cast<int>(l.size())));
                            elam.push_back(l);
                        }
                        if (need > static_cast<size_t>(max_lds)) ok = false;
                    }
                    if (!ok) continue;
                    cond_of[b] = 1;
                    sys->cond_lds_bytes = std::max(sys->cond_lds_bytes, need);
                    // interface rows: dofs of nodes outside every element, and all constraint rows
                    std::vector<int> zidx(B.block_n[b], -1);
                    for (int q = B.block_node_begin[b]; q < B.block_node_end[b]; ++q) {
                        const int node = B.block_nodes[q];
                        if (!owner.count(node))
                            for (int c = 0; c < 6; ++c) zidx[B.node_local[node] + c] = 0;
                    }
                    for (int i = B.block_nsys[b]; i < B.block_n[b]; ++i) zidx[i] = 0;
                    int nz = 0;
                    for (int i = 0; i < B.block_n[b]; ++i)
                        if (zidx[i] == 0) {
                            zidx[i] = nz++;
                            zl.push_back(i);
                        }
                    nz_of[b] = nz;
                    for (int i = 0; i < B.block_n[b]; ++i) dof_z[B.block_vec_off[b] + i] = zidx[i];
                    {   // tree schedule of the interface system, if its graph is a forest
                        TreeSchedule T;
                        if (build_interface_tree(B, b, owner, elam, zidx, T)) {
                            const int vb = static_cast<int>(tv_size.size());
                            tree_block[6 * b + 0] = vb;
                            tree_block[6 * b + 1] = static_cast<int>(T.size.size());
                            for (size_t v = 0; v < T.size.size(); ++v) {
                                tv_size.push_back(T.size[v]);
                                tv_row_off.push_back(static_cast<int>(tv_rows.size()));
                                tv_rows.insert(tv_rows.end(), T.rows.begin() + T.row_off[v], T.rows.begin() + T.row_off[v + 1]);
                                tv_parent.push_back(T.parent[v] >= 0 ? vb + T.parent[v] : -1);
                                tv_child_off.push_back(static_cast<int>(tv_children.size()));
                                for (int k = T.child_off[v]; k < T.child_off[v + 1]; ++k) tv_children.push_back(vb + T.children[k]);
                            }
                            auto put = [&](const std::vector<int>& ids, const std::vector<int>& off, std::vector<int>& gids,
                                           std::vector<int>& goff, int slot) {
                                tree_block[6 * b + slot] = static_cast<int>(goff.size());
                                tree_block[6 * b + slot + 1] = static_cast<int>(off.size()) - 1;
                                const int base = static_cast<int>(gids.size());
                                for (int o : off) goff.push_back(base + o);
                                for (int v : ids) gids.push_back(vb + v);
                            };
                            put(T.elim, T.elim_lvl, t_elim, t_elim_lvl, 2);
                            put(T.back, T.back_lvl, t_back, t_back_lvl, 4);
                        }
                    }
                    Z_off[b] = Z_tot;
                    Z_tot += static_cast<int64_t>(nz) * nz;
                    int idx = 0;
                    for (int k = B.block_elem_begin[b]; k < B.block_elem_end[b]; ++k, ++idx) {
                        const int e = B.block_elems[k];
                        ce_elem.push_back(e);
                        ce_block.push_back(b);
                        for (int i = 0; i < B.elem_N[e]; ++i)
                            for (int c = 0; c < 6; ++c) {
                                const int r = B.node_local[B.elem_nodes[B.elem_node_off[e] + i]] + c;
                                dof_ce[B.block_vec_off[b] + r] = static_cast<int>(ce_elem.size()) - 1;
                                dof_cl[B.block_vec_off[b] + r] = 6 * i + c;
                                rows.push_back(r);
                            }
                        rows_off.push_back(static_cast<int>(rows.size()));
                        for (int r : elam[idx]) {
                            lam.push_back(r);
                            lam_z.push_back(zidx[r]);
                        }
                        lam_off.push_back(static_cast<int>(lam.size()));
                        const int ne = 6 * B.elem_N[e], me = static_cast<int>(elam[idx].size());
                        {   // gather maps: the element's rows into [K_e | B_e^T], and B_e's entries
                            const int* rpb = rp_all.data() + sys->crs_rp_off[b];
                            const int* cib = col_all.data() + sys->crs_nz_off[b];
                            const int vb = B.block_vec_off[b], ce = static_cast<int>(ce_elem.size()) - 1;
                            const int lda = (ne + me + 1) | 1;
                            const int* er = rows.data() + rows_off[rows_off.size() - 2];
                            for (int i = 0; i < ne; ++i)
                                for (int p = rpb[er[i]]; p < rpb[er[i] + 1]; ++p) {
                                    const int c = cib[p];
                                    int dst = -1;
                                    if (dof_ce[vb + c] == ce) dst = i * lda + dof_cl[vb + c];
                                    for (int k = 0; k < me && dst < 0; ++k)
                                        if (elam[idx][k] == c) dst = i * lda + ne + k;
                                    if (dst >= 0) {
                                        gat_src.push_back(p);
                                        gat_dst.push_back(dst);
                                    }
                                }
                            gat_off.push_back(static_cast<int>(gat_src.size()));
                            for (int k = 0; k < me; ++k) {
                                const int row = elam[idx][k];
                                for (int p = rpb[row]; p < rpb[row + 1]; ++p)
                                    if (dof_ce[vb + cib[p]] == ce) {
                                        bt_src.push_back(p);
                                        bt_col.push_back(dof_cl[vb + cib[p]]);
                                    }
                                bt_off.push_back(static_cast<int>(bt_src.size()));
                            }
                        }
                        Y_off.push_back(Y_tot);
                        Y_tot += static_cast<int64_t>(ne) * (me + 1);
                        S_off.push_back(S_tot);
                        S_tot += static_cast<int64_t>(me) * (me + 1);
                    }
                    nce_of[b] = B.block_elem_end[b] - B.block_elem_begin[b];
                    {   // colour the block's condensed elements: no two of one colour share a constraint row
                        const int ce0 = static_cast<int>(ce_elem.size()) - nce_of[b];
                        std::vector<int> colour(nce_of[b], -1);
                        int ncol = 0;
                        for (int a = 0; a < nce_of[b]; ++a) {
                            std::set<int> used;
                            for (int c = 0; c < a; ++c) {
                                bool share = false;
                                for (int r1 : elam[a])
                                    for (int r2 : elam[c]) share |= r1 == r2;
                                if (share) used.insert(colour[c]);
                            }
                            int k = 0;
                            while (used.count(k)) ++k;
                            colour[a] = k;
                            ncol = std::max(ncol, k + 1);
                        }
                        block_color[2 * b] = static_cast<int>(color_off.size()) - 1;
                        block_color[2 * b + 1] = ncol;
                        sys->colour_sizes.resize(B.block_n.size());
                        for (int k = 0; k < ncol; ++k) {
                            for (int a = 0; a < nce_of[b]; ++a)
                                if (colour[a] == k) color_ce.push_back(ce0 + a);
                            sys->colour_sizes[b].push_back(static_cast<int>(color_ce.size()) - color_off.back());
                            color_off.push_back(static_cast<int>(color_ce.size()));
                        }
                    }
                }
                // per-block ranges (blocks in order; only condensed blocks are non-empty)
                for (size_t b = 0; b + 1 < z_off.size(); ++b) {
                    z_off[b + 1] = z_off[b] + nz_of[b];
                    ce_begin[b + 1] = ce_begin[b] + nce_of[b];
                }
                S.num_cond_elems = static_cast<int>(ce_elem.size());
                S.cond_elem = Ar.up(ce_elem);
                S.cond_elem_block = Ar.up(ce_block);
                S.cond_rows_off = Ar.up(rows_off);
                S.cond_rows = Ar.up(rows);
                S.cond_lam_off = Ar.up(lam_off);
                S.cond_lam = Ar.up(lam);
                S.cond_lam_z = Ar.up(lam_z);
                S.cond_Y_off = Ar.up(Y_off);
                S.cond_S_off = Ar.up(S_off);
                S.Ybuf = Ar.zeros<double>(static_cast<size_t>(Y_tot));
                S.Sbuf = Ar.zeros<double>(static_cast<size_t>(S_tot));
                S.block_z_off = Ar.up(z_off);
                S.cond_z = Ar.up(zl);
                S.block_ce_begin = Ar.up(ce_begin);
                S.block_Z_off = Ar.up(Z_off);
                S.Zbuf = Ar.zeros<double>(static_cast<size_t>(Z_tot));
                tv_row_off.push_back(static_cast<int>(tv_rows.size()));
                tv_child_off.push_back(static_cast<int>(tv_children.size()));
                S.block_color = Ar.up(block_color);
                S.cond_color_off = Ar.up(color_off);
                S.cond_color_ce = Ar.up(color_ce);
                S.tree_block = Ar.up(tree_block);
                S.tree_vsize = Ar.up(tv_size);
                S.tree_vrow_off = Ar.up(tv_row_off);
                S.tree_vrows = Ar.up(tv_rows);
                S.tree_vparent = Ar.up(tv_parent);
                S.tree_vchild_off = Ar.up(tv_child_off);
                S.tree_vchildren = Ar.up(tv_children);
                S.tree_elim = Ar.up(t_elim);
                S.tree_elim_lvl = Ar.up(t_elim_lvl);
                S.tree_back = Ar.up(t_back);
                S.tree_back_lvl = Ar.up(t_back_lvl);
                sys->tree_nvert.assign(B.block_n.size(), 0);
                for (size_t bb = 0; bb < B.block_n.size(); ++bb) sys->tree_nvert[bb] = tree_block[6 * bb + 1];
                S.dof_cond_elem = Ar.up(dof_ce);
                S.dof_cond_local = Ar.up(dof_cl);
                S.cond_gat_off = Ar.up(gat_off);
                S.cond_gat_src = Ar.up(gat_src);
                S.cond_gat_dst = Ar.up(gat_dst);
                S.cond_bt_off = Ar.up(bt_off);
                S.cond_bt_src = Ar.up(bt_src);
                S.cond_bt_col = Ar.up(bt_col);
                S.dof_zidx = Ar.up(dof_z);
                S.zrhs = Ar.zeros<double>(static_cast<size_t>(z_off.back()));
                for (size_t k = 0; k < sys->large_blocks.size(); ++k) {
                    const int b = sys->large_blocks[k];
                    sys->large_cond.push_back(cond_of[b]);
                    sys->large_nz.push_back(nz_of[b]);
                    sys->large_Z_off.push_back(Z_off[b]);
                }
            }
            {
                // dense column-major matrices of the large blocks that do not condense (library LU)
                std::vector<int64_t> big_off(B.block_n.size(), 0);
                int64_t dense = 0;
                for (size_t k = 0; k < sys->large_blocks.size(); ++k) {
                    sys->large_off[k] = dense;
                    if (sys->large_cond.empty() || !sys->large_cond[k]) {
                        big_off[sys->large_blocks[k]] = dense;
                        dense += static_cast<int64_t>(sys->large_n[k]) * sys->large_n[k];
                    }
                }
                S.block_big_off = Ar.up(big_off);
                S.Abig = dense > 0 ? Ar.zeros<double>(static_cast<size_t>(dense)) : nullptr;
                S.Abig32 = (S.precision != KYN_TANGENT_FP64 && dense > 0) ? Ar.zeros<float>(static_cast<size_t>(dense))
                                                                          : nullptr;
            }
            if (!sys->large_blocks.empty()) {
                const int nbig = *std::max_element(sys->large_n.begin(), sys->large_n.end());
                sys->ipiv = Ar.zeros<rocblas_int>(static_cast<size_t>(nbig));
                sys->info = Ar.zeros<rocblas_int>(1);
                sys->rvec = Ar.zeros<double>(static_cast<size_t>(nbig));
                sys->bvec = Ar.zeros<double>(static_cast<size_t>(nbig));
                sys->r32 = Ar.zeros<float>(static_cast<size_t>(nbig));
                sys->rfull = Ar.zeros<double>(static_cast<size_t>(B.vec_total));
                sys->dfull = Ar.zeros<double>(static_cast<size_t>(B.vec_total));
                if (rocblas_create_handle(&sys->blas) != rocblas_status_success)
                    throw KynError(KYN_ERR_HIP, "rocblas_create_handle failed");
            }
        }
        S.block_vec_off = Ar.up(B.block_vec_off);
        S.node_block = Ar.up(B.node_block);
        S.node_local = Ar.up(B.node_local);
        S.node_pin = Ar.up(B.node_pin);
        S.block_npin = Ar.up(B.block_npin);
        S.block_nopiv = Ar.up(B.block_nopiv);
        S.elem_N = Ar.up(B.elem_N);
        S.elem_Q = Ar.up(B.elem_Q);
        S.elem_node_off = Ar.up(B.elem_node_off);
        S.elem_nodes = Ar.up(B.elem_nodes);
        S.elem_qp_off = Ar.up(B.elem_qp_off);
        S.elem_shape_off = Ar.up(B.elem_shape_off);
        S.elem_mat_off = Ar.up(B.elem_mat_off);
        S.qp_weight = Ar.up(B.qp_weight);
        S.qp_jac = Ar.up(B.qp_jac);
        S.qp_Mstar = Ar.up(B.qp_Mstar);
        S.qp_Cstar = Ar.up(B.qp_Cstar);
        S.qp_x0p = Ar.up(B.qp_x0p);
        S.qp_r0 = Ar.up(B.qp_r0);
        S.shape_interp = Ar.up(B.shape_interp);
        S.shape_deriv = Ar.up(B.shape_deriv);
        S.node_FX = Ar.zeros<double>(B.elem_nodes.size() * 6);
        S.elem_res = Ar.zeros<double>(B.elem_nodes.size() * 6);
        // element matrices and dense KKT blocks in HBM exist only in the plain form (the fused
        // form keeps them on chip); kyn_debug_assemble allocates them on demand
        S.elem_mat = sys->exec_mode == KYN_EXEC_PLAIN ? Ar.zeros<double>(static_cast<size_t>(B.mat_total)) : nullptr;
        sys->mat_total = B.mat_total;
        sys->A_total = B.A_total;
        S.block_mass_begin = Ar.up(B.block_mass_begin);
        S.block_mass_end = Ar.up(B.block_mass_end);
        S.block_masses = Ar.up(B.block_masses);
        S.block_spring_begin = Ar.up(B.block_spring_begin);
        S.block_spring_end = Ar.up(B.block_spring_end);
        S.block_springs = Ar.up(B.block_springs);
        S.mass_node = Ar.up(B.mass_node);
        S.mass_Mstar = Ar.up(B.mass_Mstar);
        S.mass_r0 = Ar.up(B.mass_r0);
        S.mass_res = Ar.zeros<double>(B.mass_node.size() * 6);
        S.mass_mat = Ar.zeros<double>(B.mass_node.size() * 36);
        S.spring_n1 = Ar.up(B.spring_n1);
        S.spring_n2 = Ar.up(B.spring_n2);
        S.spring_x0 = Ar.up(B.spring_x0);
        S.spring_lref = Ar.up(B.spring_lref);
        S.spring_k = Ar.up(B.spring_k);
        S.spring_res = Ar.zeros<double>(B.spring_n1.size() * 6);
        S.spring_mat = Ar.zeros<double>(B.spring_n1.size() * 9);
        const size_t nc = B.con_type.size();
        S.con_type = Ar.up(B.con_type);
        S.con_base = Ar.up(B.con_base);
        S.con_target = Ar.up(B.con_target);
        S.con_rows = Ar.up(B.con_rows);
        S.con_model_row = Ar.up(B.con_model_row);
        S.con_local = Ar.up(B.con_local);
        S.con_X0 = Ar.up(B.con_X0);
        S.con_axes = Ar.up(B.con_axes);
        S.con_input = Ar.up(B.con_input);
        S.con_control = Ar.zeros<double>(nc);
        S.con_lambda = Ar.zeros<double>(nc * 6);
        S.con_res = Ar.zeros<double>(nc * 6);
        S.con_sys_res = Ar.zeros<double>(nc * 6);
        S.con_output = Ar.zeros<double>(nc * 3);
        S.con_tgrad = Ar.zeros<double>(nc * 36);
        S.con_bgrad = Ar.zeros<double>(nc * 36);
        S.con_tgradT = Ar.zeros<double>(nc * 36);
        S.con_bgradT = Ar.zeros<double>(nc * 36);
        S.con_tlres = Ar.zeros<double>(nc * 6);
        S.con_blres = Ar.zeros<double>(nc * 6);
        const size_t tt = S.precision ? sizeof(float) : sizeof(double);
        S.A = sys->exec_mode == KYN_EXEC_PLAIN ? Ar.zeros<double>(static_cast<size_t>(B.A_total)) : nullptr;
        S.Acopy = (S.refine_sweeps > 0 && sys->exec_mode == KYN_EXEC_FUSED)
                      ? static_cast<void*>(Ar.zeros<char>(static_cast<size_t>(B.C_total) * tt))
                      : nullptr;
        S.rhs = Ar.zeros<double>(B.vec_total);
        S.xsol = Ar.zeros<double>(B.vec_total);
        S.rhs_raw = Ar.zeros<double>(B.vec_total);
        S.err_hist = Ar.zeros<double>(num_models * static_cast<size_t>(S.p.max_iter));
        S.model_iters = Ar.zeros<int>(num_models);
        S.model_status = Ar.zeros<int>(num_models);
        S.model_active = Ar.zeros<int>(num_models);
        S.model_iter_total = Ar.zeros<int>(num_models);
        S.model_fail_total = Ar.zeros<int>(num_models);
        S.any_active = Ar.zeros<int>(1);
        S.sched_slot = Ar.zeros<int>(1);
        S.dev_error = Ar.zeros<int>(1);
        S.lu_wait_limit = kyn::kLuWaitLimitDefault;
        S.sched_steps = 0;

        // ---- host copies of the layout
        sys->num_models = num_models;
        sys->num_nodes = n;
        sys->num_cons = nc;
        sys->num_elems = B.elem_N.size();
        sys->num_blocks = B.block_n.size();
        sys->model_node_begin = B.model_node_begin;
        sys->model_node_end = B.model_node_end;
        sys->model_con_begin = B.model_con_begin;
        sys->model_con_end = B.model_con_end;
        sys->model_ndofs = B.model_ndofs;
        sys->model_nsys = B.model_nsys;
        sys->elem_N = B.elem_N;
        sys->elem_node_off = B.elem_node_off;
        sys->block_n = B.block_n;
        sys->block_vec_off = B.block_vec_off;
        sys->block_model = B.block_model;
        sys->node_block = B.node_block;
        sys->node_local = B.node_local;
        sys->con_local = B.con_local;
        sys->con_rows = B.con_rows;
        sys->con_model_row = B.con_model_row;
        sys->con_block = B.con_block;
        sys->model_gdof_to_block_vec = B.gdof_maps;

        hip_check(hipStreamCreateWithFlags(&sys->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (sys->blas && rocblas_set_stream(sys->blas, sys->stream) != rocblas_status_success)
            throw KynError(KYN_ERR_HIP, "rocblas_set_stream failed");
        hip_check(hipEventCreate(&sys->ev0), "event");
        hip_check(hipEventCreate(&sys->ev1), "event");
        hip_check(hipEventCreate(&sys->ek0), "event");
        hip_check(hipEventCreate(&sys->ek1), "event");
        // UpdateGlobalPosition at setup (copy_nodes_to_state.hpp:52-58) via the epilogue's x
        {
            std::vector<double> x(n * 7);
            for (size_t i = 0; i < n; ++i) {
                const double* q = &B.q[i * 7];
                const double* x0 = &B.x0[i * 7];
                x[i * 7 + 0] = x0[0] + q[0];
                x[i * 7 + 1] = x0[1] + q[1];
                x[i * 7 + 2] = x0[2] + q[2];
                const auto r = qcompose({q[3], q[4], q[5], q[6]}, {x0[3], x0[4], x0[5], x0[6]});
                for (int c = 0; c < 4; ++c) x[i * 7 + 3 + c] = r[c];
            }
            hip_check(hipMemcpy(S.x, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
        }
        result = sys.release();
        return KYN_OK;
    });
    (void)rc;
    return result;
}

void kyn_system_destroy(kyn_system s) { delete s; }
size_t kyn_system_num_models(kyn_system s) { return s ? s->num_models : 0; }
size_t kyn_system_num_nodes(kyn_system s) { return s ? s->num_nodes : 0; }
size_t kyn_system_num_constraints(kyn_system s) { return s ? s->num_cons : 0; }
size_t kyn_system_num_dofs(kyn_system s, size_t model) {
    return (s && model < s->num_models) ? static_cast<size_t>(s->model_ndofs[model]) : 0;
}
size_t kyn_system_node_offset(kyn_system s, size_t model) {
    return (s && model < s->num_models) ? static_cast<size_t>(s->model_node_begin[model]) : 0;
}
size_t kyn_system_constraint_offset(kyn_system s, size_t model) {
    return (s && model < s->num_models) ? static_cast<size_t>(s->model_con_begin[model]) : 0;
}

int kyn_step(kyn_system s) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        return run_step(s);
    });
}

int kyn_step_many(kyn_system s, size_t num_steps, int use_graph) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        if (num_steps == 0) return KYN_OK;
        auto& S = s->S;
        if (s->exec_mode != KYN_EXEC_FUSED) {
            int rc = KYN_OK;
            float kms = 0.f, tms = 0.f;
            for (size_t k = 0; k < num_steps; ++k) {
                if (run_step(s) != KYN_OK) rc = KYN_NOT_CONVERGED;
                kms += s->last_kernel_ms;
                tms += s->last_step_ms;
            }
            s->last_kernel_ms = kms / static_cast<float>(num_steps);
            s->last_step_ms = tms;
            return rc;
        }
        if (use_graph && !s->graph_ready) {
            if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
            if (s->graph) (void)hipGraphDestroy(s->graph);
            s->graph_exec = nullptr;
            s->graph = nullptr;
            hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            kyn::launch_fused_step(S, s->Nm, s->Qm, s->nm, s->stream);
            kyn::launch_advance_slot(S, s->stream);
            hip_check(hipStreamEndCapture(s->stream, &s->graph), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            s->graph_ready = true;
        }
        std::vector<hipEvent_t> ev(2 * num_steps);
        for (auto& e : ev) hip_check(hipEventCreate(&e), "event");
        hip_check(hipEventRecord(s->ev0, s->stream), "event");
        for (size_t k = 0; k < num_steps; ++k) {
            hip_check(hipEventRecord(ev[2 * k], s->stream), "event");
            if (use_graph) {
                hip_check(hipGraphLaunch(s->graph_exec, s->stream), "hipGraphLaunch");
                hip_check(hipEventRecord(ev[2 * k + 1], s->stream), "event");
            } else {
                kyn::launch_fused_step(S, s->Nm, s->Qm, s->nm, s->stream);
                hip_check(hipGetLastError(), "k_fused_step launch");
                hip_check(hipEventRecord(ev[2 * k + 1], s->stream), "event");
                kyn::launch_advance_slot(S, s->stream);
            }
        }
        hip_check(hipEventRecord(s->ev1, s->stream), "event");
        hip_check(hipStreamSynchronize(s->stream), "sync");
        float total = 0.f, kern = 0.f;
        hip_check(hipEventElapsedTime(&total, s->ev0, s->ev1), "elapsed");
        for (size_t k = 0; k < num_steps; ++k) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]), "elapsed");
            kern += ms;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        s->last_step_ms = total;
        s->last_kernel_ms = kern / static_cast<float>(num_steps);
        check_device_errors(s);
        std::vector<int> fails(s->num_models);
        hip_check(hipMemcpy(fails.data(), S.model_fail_total, fails.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        for (int f : fails)
            if (f != 0) return KYN_NOT_CONVERGED;
        return KYN_OK;
    });
}

int kyn_get_totals(kyn_system s, int* iterations, int* failures) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        if (iterations)
            hip_check(hipMemcpy(iterations, s->S.model_iter_total, s->num_models * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        if (failures)
            hip_check(hipMemcpy(failures, s->S.model_fail_total, s->num_models * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        return KYN_OK;
    });
}

int kyn_get_step_status(kyn_system s, int* status, int* iterations) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        if (status)
            hip_check(hipMemcpy(status, s->S.model_status, s->num_models * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        if (iterations)
            hip_check(hipMemcpy(iterations, s->S.model_iters, s->num_models * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        return KYN_OK;
    });
}

size_t kyn_get_convergence_history(kyn_system s, size_t model, double* err, size_t cap) {
    size_t count = 0;
    guarded([&] {
        if (!s || model >= s->num_models) throw KynError(KYN_ERR_INVALID, "bad model");
        int it = 0;
        hip_check(hipMemcpy(&it, s->S.model_iters + model, sizeof(int), hipMemcpyDeviceToHost), "D2H");
        count = static_cast<size_t>(it);
        if (err && cap) {
            const size_t k = std::min(cap, count);
            hip_check(hipMemcpy(err, s->S.err_hist + model * static_cast<size_t>(s->S.p.max_iter), k * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "D2H");
        }
        return KYN_OK;
    });
    return count;
}

double kyn_last_step_ms(kyn_system s) { return s ? static_cast<double>(s->last_step_ms) : 0.; }
double kyn_last_kernel_ms(kyn_system s) { return s ? static_cast<double>(s->last_kernel_ms) : 0.; }

int kyn_set_constraint_displacement(kyn_system s, size_t id, const double disp[7]) {
    return kyn_set_constraint_displacements(s, 1, &id, disp);
}
int kyn_set_constraint_displacements(kyn_system s, size_t count, const size_t* ids, const double* disp) {
    return guarded([&] {
        if (!s || !ids || !disp) throw KynError(KYN_ERR_INVALID, "null argument");
        for (size_t k = 0; k < count; ++k) {
            if (ids[k] >= s->num_cons) throw KynError(KYN_ERR_INVALID, "constraint id out of range");
            hip_check(hipMemcpyAsync(s->S.con_input + ids[k] * 7, disp + 7 * k, 7 * sizeof(double),
                                     hipMemcpyHostToDevice, s->stream),
                      "H2D");
        }
        hip_check(hipStreamSynchronize(s->stream), "sync");
        return KYN_OK;
    });
}
int kyn_set_constraint_control(kyn_system s, size_t id, double value) {
    return guarded([&] {
        if (!s || id >= s->num_cons) throw KynError(KYN_ERR_INVALID, "constraint id out of range");
        hip_check(hipMemcpy(s->S.con_control + id, &value, sizeof(double), hipMemcpyHostToDevice), "H2D");
        return KYN_OK;
    });
}
int kyn_set_nodal_loads(kyn_system s, const double* f) {
    return guarded([&] {
        if (!s || !f) throw KynError(KYN_ERR_INVALID, "null argument");
        hip_check(hipMemcpy(s->S.f, f, s->num_nodes * 6 * sizeof(double), hipMemcpyHostToDevice), "H2D");
        return KYN_OK;
    });
}
int kyn_set_element_node_loads(kyn_system s, size_t element, const double* fx) {
    return guarded([&] {
        if (!s || !fx || element >= s->num_elems) throw KynError(KYN_ERR_INVALID, "bad element");
        hip_check(hipMemcpy(s->S.node_FX + static_cast<size_t>(s->elem_node_off[element]) * 6, fx,
                            static_cast<size_t>(s->elem_N[element]) * 6 * sizeof(double), hipMemcpyHostToDevice),
                  "H2D");
        return KYN_OK;
    });
}
int kyn_set_input_schedule(kyn_system s, size_t num_steps, const double* cdisp, size_t num_loaded,
                           const size_t* load_nodes, const double* loads) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        auto& S = s->S;
        auto& Ar = s->arena;
        S.sched_steps = static_cast<int>(num_steps);
        S.sched_cdisp = nullptr;
        S.sched_loads = nullptr;
        S.sched_num_loaded = 0;
        if (cdisp && s->num_cons > 0)
            S.sched_cdisp = Ar.up(std::vector<double>(cdisp, cdisp + num_steps * s->num_cons * 7));
        std::vector<int> begin(s->num_models + 1, 0);
        if (loads && num_loaded > 0) {
            std::vector<size_t> order(num_loaded);
            std::iota(order.begin(), order.end(), size_t{0});
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return load_nodes[a] < load_nodes[b]; });
            std::vector<int> nodes(num_loaded);
            std::vector<double> vals(num_steps * num_loaded * 6);
            for (size_t k = 0; k < num_loaded; ++k) {
                if (load_nodes[order[k]] >= s->num_nodes) throw KynError(KYN_ERR_INVALID, "load node out of range");
                nodes[k] = static_cast<int>(load_nodes[order[k]]);
                for (size_t st = 0; st < num_steps; ++st)
                    for (int c = 0; c < 6; ++c)
                        vals[(st * num_loaded + k) * 6 + c] = loads[(st * num_loaded + order[k]) * 6 + c];
            }
            size_t k = 0;
            for (size_t m = 0; m < s->num_models; ++m) {
                begin[m] = static_cast<int>(k);
                while (k < num_loaded && nodes[k] < s->model_node_end[m]) ++k;
            }
            begin[s->num_models] = static_cast<int>(num_loaded);
            S.sched_load_nodes = Ar.up(nodes);
            S.sched_loads = Ar.up(vals);
            S.sched_num_loaded = static_cast<int>(num_loaded);
        }
        hip_check(hipMemcpy(const_cast<int*>(S.model_load_begin), begin.data(), begin.size() * sizeof(int),
                            hipMemcpyHostToDevice),
                  "H2D");
        const int zero = 0;
        hip_check(hipMemcpy(S.sched_slot, &zero, sizeof(int), hipMemcpyHostToDevice), "H2D");
        s->graph_ready = false;
        return KYN_OK;
    });
}

int kyn_get_state(kyn_system s, int field, double* out) {
    return guarded([&] {
        if (!s || !out) throw KynError(KYN_ERR_INVALID, "null argument");
        if (field == KYN_TANGENT) {
            std::vector<double> tr(s->num_nodes * 9);
            hip_check(hipMemcpy(tr.data(), s->S.tan_rot, tr.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
            for (size_t i = 0; i < s->num_nodes; ++i) {
                double* T = out + i * 36;
                std::fill(T, T + 36, 0.);
                for (int k = 0; k < 3; ++k) T[k * 6 + k] = 1.;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) T[(3 + r) * 6 + 3 + c] = tr[i * 9 + r * 3 + c];
            }
            return KYN_OK;
        }
        int w = 0;
        double* p = field_ptr(s, field, w);
        if (!p) throw KynError(KYN_ERR_INVALID, "unknown field");
        hip_check(hipMemcpy(out, p, s->num_nodes * w * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        return KYN_OK;
    });
}
int kyn_set_state(kyn_system s, int field, const double* in) {
    return guarded([&] {
        if (!s || !in) throw KynError(KYN_ERR_INVALID, "null argument");
        int w = 0;
        double* p = field_ptr(s, field, w);
        if (!p) throw KynError(KYN_ERR_INVALID, "unknown field");
        hip_check(hipMemcpy(p, in, s->num_nodes * w * sizeof(double), hipMemcpyHostToDevice), "H2D");
        return KYN_OK;
    });
}
int kyn_get_constraint_values(kyn_system s, size_t id, int which, double* out) {
    return guarded([&] {
        if (!s || !out || id >= s->num_cons) throw KynError(KYN_ERR_INVALID, "bad constraint");
        const double* src = which == 0 ? s->S.con_lambda + id * 6
                            : which == 1 ? s->S.con_res + id * 6
                            : which == 2 ? s->S.con_output + id * 3
                                         : nullptr;
        if (!src) throw KynError(KYN_ERR_INVALID, "bad selector");
        hip_check(hipMemcpy(out, src, (which == 2 ? 3 : 6) * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        return KYN_OK;
    });
}
static int get_vec(kyn_system s, size_t model, double* out, const double* dev) {
    return guarded([&] {
        if (!s || !out || model >= s->num_models) throw KynError(KYN_ERR_INVALID, "bad model");
        std::vector<double> all(std::accumulate(s->block_n.begin(), s->block_n.end(), size_t{0}));
        hip_check(hipMemcpy(all.data(), dev, all.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        const auto& g = s->model_gdof_to_block_vec[model];
        for (size_t i = 0; i < g.size(); ++i) out[i] = g[i] >= 0 ? all[static_cast<size_t>(g[i])] : 0.;
        return KYN_OK;
    });
}
int kyn_get_rhs(kyn_system s, size_t model, double* b) { return get_vec(s, model, b, s ? s->S.rhs_raw : nullptr); }
int kyn_get_solution(kyn_system s, size_t model, double* x) { return get_vec(s, model, x, s ? s->S.xsol : nullptr); }

// CopyStateData (copy_state_data.hpp:20-34): x, q_delta, q_prev, q, v, vd, a, f
int kyn_get_motion(kyn_system s, double* out) {
    return guarded([&] {
        if (!s || !out) throw KynError(KYN_ERR_INVALID, "null argument");
        if (!s->motion) s->motion = s->arena.zeros<double>(s->num_nodes * 26);
        kyn::launch_pack_motion(s->S, s->motion, s->stream);
        hip_check(hipGetLastError(), "k_pack_motion launch");
        hip_check(hipMemcpyAsync(out, s->motion, s->num_nodes * 26 * sizeof(double), hipMemcpyDeviceToHost, s->stream),
                  "D2H");
        hip_check(hipStreamSynchronize(s->stream), "sync");
        return KYN_OK;
    });
}

size_t kyn_model_num_nodes(kyn_model m) { return m ? m->nodes.size() : 0; }
int kyn_model_get_node(kyn_model m, size_t id, double x0[7], double u[7], double v[6], double vd[6], double* s) {
    if (!m || id >= m->nodes.size()) return KYN_ERR_INVALID;
    const auto& n = m->nodes[id];
    if (x0) std::copy(n.x0.begin(), n.x0.end(), x0);
    if (u) std::copy(n.u.begin(), n.u.end(), u);
    if (v) std::copy(n.v.begin(), n.v.end(), v);
    if (vd) std::copy(n.vd.begin(), n.vd.end(), vd);
    if (s) *s = n.s;
    return KYN_OK;
}
int kyn_model_beam_element_info(kyn_model m, size_t elem, size_t* num_nodes, size_t* num_sections, size_t* num_qps) {
    if (!m || elem >= m->beams.size()) return KYN_ERR_INVALID;
    const auto& b = m->beams[elem];
    if (num_nodes) *num_nodes = b.nodes.size();
    if (num_sections) *num_sections = b.sections.size();
    if (num_qps) *num_qps = b.quad.size();
    return KYN_OK;
}
int kyn_model_get_beam_element(kyn_model m, size_t elem, size_t* node_ids, double* section_s, double* section_mass,
                               double* section_stiffness, double* qp_xi_w) {
    if (!m || elem >= m->beams.size()) return KYN_ERR_INVALID;
    const auto& b = m->beams[elem];
    if (node_ids) std::copy(b.nodes.begin(), b.nodes.end(), node_ids);
    for (size_t k = 0; k < b.sections.size(); ++k) {
        if (section_s) section_s[k] = b.sections[k].s;
        if (section_mass) std::copy(b.sections[k].M.begin(), b.sections[k].M.end(), section_mass + 36 * k);
        if (section_stiffness) std::copy(b.sections[k].C.begin(), b.sections[k].C.end(), section_stiffness + 36 * k);
    }
    if (qp_xi_w)
        for (size_t q = 0; q < b.quad.size(); ++q) {
            qp_xi_w[2 * q] = b.quad[q][0];
            qp_xi_w[2 * q + 1] = b.quad[q][1];
        }
    return KYN_OK;
}

int kyn_save_state(kyn_system s) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        const size_t n = s->num_nodes;
        s->saved.assign(n * (7 * 3 + 6 * 5) + 1, 0.);
        double* o = s->saved.data();
        const int fields[] = {KYN_X, KYN_Q_DELTA, KYN_Q_PREV, KYN_Q, KYN_V, KYN_VD, KYN_A, KYN_F};
        for (int f : fields) {
            int w = 0;
            double* p = field_ptr(s, f, w);
            hip_check(hipMemcpy(o, p, n * w * sizeof(double), hipMemcpyDeviceToHost), "D2H");
            o += n * w;
        }
        int slot = 0;
        hip_check(hipMemcpy(&slot, s->S.sched_slot, sizeof(int), hipMemcpyDeviceToHost), "D2H");
        *o = static_cast<double>(slot);
        s->has_saved = true;
        return KYN_OK;
    });
}
int kyn_restore_state(kyn_system s) {
    return guarded([&] {
        if (!s || !s->has_saved) throw KynError(KYN_ERR_INVALID, "no saved state");
        const size_t n = s->num_nodes;
        const double* o = s->saved.data();
        const int fields[] = {KYN_X, KYN_Q_DELTA, KYN_Q_PREV, KYN_Q, KYN_V, KYN_VD, KYN_A, KYN_F};
        for (int f : fields) {
            int w = 0;
            double* p = field_ptr(s, f, w);
            hip_check(hipMemcpy(p, o, n * w * sizeof(double), hipMemcpyHostToDevice), "H2D");
            o += n * w;
        }
        const int slot = static_cast<int>(*o);
        hip_check(hipMemcpy(s->S.sched_slot, &slot, sizeof(int), hipMemcpyHostToDevice), "H2D");
        return KYN_OK;
    });
}
// WriteStateToFile / ReadStateFromFile byte image (write_state_to_file.hpp:19-58)
size_t kyn_restart_size(kyn_system s, size_t model) {
    if (!s || model >= s->num_models) return 0;
    const size_t n = static_cast<size_t>(s->model_node_end[model] - s->model_node_begin[model]);
    return sizeof(size_t) + n * (7 * 4 + 6 * 5) * sizeof(double);
}
int kyn_write_restart(kyn_system s, size_t model, unsigned char* out) {
    return guarded([&] {
        if (!s || !out || model >= s->num_models) throw KynError(KYN_ERR_INVALID, "bad model");
        size_t n = static_cast<size_t>(s->model_node_end[model] - s->model_node_begin[model]);
        const size_t n0 = static_cast<size_t>(s->model_node_begin[model]);
        std::memcpy(out, &n, sizeof(size_t));
        size_t off = sizeof(size_t);
        const int fields[] = {KYN_X0, KYN_X, KYN_Q_DELTA, KYN_Q_PREV, KYN_Q, KYN_V, KYN_VD, KYN_A, KYN_F};
        for (int f : fields) {
            int w = 0;
            double* p = field_ptr(s, f, w);
            hip_check(hipMemcpy(out + off, p + n0 * w, n * w * sizeof(double), hipMemcpyDeviceToHost), "D2H");
            off += n * w * sizeof(double);
        }
        return KYN_OK;
    });
}
int kyn_read_restart(kyn_system s, size_t model, const unsigned char* in, size_t size) {
    return guarded([&] {
        if (!s || !in || model >= s->num_models) throw KynError(KYN_ERR_INVALID, "bad model");
        if (size < sizeof(size_t)) throw KynError(KYN_ERR_INVALID, "restart image too short for its node count");
        size_t n = 0;
        std::memcpy(&n, in, sizeof(size_t));
        const size_t nm = static_cast<size_t>(s->model_node_end[model] - s->model_node_begin[model]);
        // read_state_from_file.hpp:28-30 (std::length_error there)
        if (n != nm) throw KynError(KYN_ERR_INVALID, "Number of system nodes in file is not the same as in model");
        if (size < kyn_restart_size(s, model)) throw KynError(KYN_ERR_INVALID, "restart image too short");
        const size_t n0 = static_cast<size_t>(s->model_node_begin[model]);
        size_t off = sizeof(size_t);
        const int fields[] = {KYN_X0, KYN_X, KYN_Q_DELTA, KYN_Q_PREV, KYN_Q, KYN_V, KYN_VD, KYN_A, KYN_F};
        for (int f : fields) {
            int w = 0;
            double* p = field_ptr(s, f, w);
            hip_check(hipMemcpy(p + n0 * w, in + off, n * w * sizeof(double), hipMemcpyHostToDevice), "H2D");
            off += n * w * sizeof(double);
        }
        return KYN_OK;
    });
}

int kyn_debug_set_contraction_mask(kyn_system s, unsigned mask) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        s->S.contract_mfma_mask = mask;
        s->graph_ready = false;  // the captured step holds the kernel arguments
        return KYN_OK;
    });
}

int kyn_debug_set_wait_limit(kyn_system s, int polls) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        s->S.lu_wait_limit = polls;
        s->graph_ready = false;  // the captured step holds the kernel arguments
        return KYN_OK;
    });
}

int kyn_debug_assemble(kyn_system s) {
    return guarded([&] {
        if (!s) throw KynError(KYN_ERR_INVALID, "null system");
        if (!s->S.elem_mat) s->S.elem_mat = s->arena.zeros<double>(static_cast<size_t>(s->mat_total));
        if (!s->S.A) s->S.A = s->arena.zeros<double>(static_cast<size_t>(s->A_total));
        return run_step(s, false);
    });
}
size_t kyn_debug_get_crs(kyn_system s, size_t model, size_t* row_ptrs, size_t* col_inds, double* values, size_t cap) {
    size_t nnz = 0;
    const int rc = guarded([&] {
        if (!s || model >= s->num_models) throw KynError(KYN_ERR_INVALID, "bad model");
        const auto& g = s->model_gdof_to_block_vec[model];
        int b = -1;
        for (int v : g) {
            if (v < 0) continue;
            const int blk = static_cast<int>(std::upper_bound(s->block_vec_off.begin(), s->block_vec_off.end(), v) -
                                             s->block_vec_off.begin()) - 1;
            if (b >= 0 && blk != b) throw KynError(KYN_ERR_INVALID, "the model's dofs span several blocks");
            b = blk;
        }
        if (b < 0 || std::find(s->large_blocks.begin(), s->large_blocks.end(), b) == s->large_blocks.end())
            throw KynError(KYN_ERR_INVALID, "the model's dofs are not one large block (no CRS matrix)");
        const int n = s->block_n[b], vb = s->block_vec_off[b];
        std::vector<int> rp(static_cast<size_t>(n) + 1);
        hip_check(hipMemcpy(rp.data(), s->S.crs_rp + s->crs_rp_off[b], rp.size() * sizeof(int), hipMemcpyDeviceToHost),
                  "D2H");
        std::vector<int> ci(static_cast<size_t>(rp[n]));
        std::vector<double> val(ci.size());
        hip_check(hipMemcpy(ci.data(), s->S.crs_col + s->crs_nz_off[b], ci.size() * sizeof(int), hipMemcpyDeviceToHost),
                  "D2H");
        hip_check(hipMemcpy(val.data(), s->S.crs_val + s->crs_nz_off[b], val.size() * sizeof(double),
                            hipMemcpyDeviceToHost),
                  "D2H");
        std::vector<int> gdof_of(static_cast<size_t>(n), -1);  // block-local dof -> reference dof
        for (size_t i = 0; i < g.size(); ++i)
            if (g[i] >= 0) gdof_of[g[i] - vb] = static_cast<int>(i);
        std::vector<std::vector<std::pair<size_t, double>>> rows(g.size());
        for (size_t i = 0; i < g.size(); ++i) {
            if (g[i] < 0) continue;
            const int lr = g[i] - vb;
            for (int p = rp[lr]; p < rp[lr + 1]; ++p)
                if (gdof_of[ci[p]] >= 0) rows[i].emplace_back(static_cast<size_t>(gdof_of[ci[p]]), val[p]);
            std::sort(rows[i].begin(), rows[i].end());
            nnz += rows[i].size();
        }
        if (cap < nnz || !row_ptrs || !col_inds || !values) return KYN_OK;
        size_t k = 0;
        row_ptrs[0] = 0;
        for (size_t i = 0; i < g.size(); ++i) {
            for (const auto& e : rows[i]) {
                col_inds[k] = e.first;
                values[k++] = e.second;
            }
            row_ptrs[i + 1] = k;
        }
        return KYN_OK;
    });
    return rc == KYN_OK ? nnz : 0;
}

int kyn_get_element_terms(kyn_system s, size_t element, double* residual, double* system_matrix) {
    return guarded([&] {
        if (!s || element >= s->num_elems) throw KynError(KYN_ERR_INVALID, "bad element");
        if (!s->S.elem_mat) throw KynError(KYN_ERR_INVALID, "no element terms: run kyn_debug_assemble or the plain form first");
        const size_t N = static_cast<size_t>(s->elem_N[element]);
        const size_t off = static_cast<size_t>(s->elem_node_off[element]);
        if (residual)
            hip_check(hipMemcpy(residual, s->S.elem_res + off * 6, N * 6 * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        if (system_matrix) {
            int64_t moff = 0;
            hip_check(hipMemcpy(&moff, s->S.elem_mat_off + element, sizeof(int64_t), hipMemcpyDeviceToHost), "D2H");
            hip_check(hipMemcpy(system_matrix, s->S.elem_mat + moff, N * N * 36 * sizeof(double), hipMemcpyDeviceToHost),
                      "D2H");
        }
        return KYN_OK;
    });
}

int kyn_element_terms_noalloc(kyn_system s, size_t element, double* residual, double* system_matrix) {
    return guarded([&] {
        if (!s || element >= s->num_elems) throw KynError(KYN_ERR_INVALID, "bad element");
        const int N = s->elem_N[element];
        if (N > kyn::kMaxElemNodes)
            throw KynError(KYN_ERR_INVALID, "element has " + std::to_string(N) + " nodes; the no-allocation element "
                                            "terms hold at most " + std::to_string(kyn::kMaxElemNodes));
        int cur = -1;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != s->device)
            throw KynError(KYN_ERR_INVALID, "make the system's device (" + std::to_string(s->device) + ") current first");
        double* scr = nullptr;
        hip_check(kyn::element_terms_scratch(s->device, &scr), "hipGetSymbolAddress(scratch)");
        double* res_d = scr;
        double* mat_d = scr + kyn::kMaxElemNodes * 6;
        kyn::launch_element_terms(s->S, s->Nm, s->Qm, static_cast<int>(element), res_d, mat_d, s->stream);
        hip_check(hipGetLastError(), "k_element_terms launch");
        const size_t n = static_cast<size_t>(N);
        if (residual)
            hip_check(hipMemcpyAsync(residual, res_d, n * 6 * sizeof(double), hipMemcpyDefault, s->stream), "copy");
        if (system_matrix)
            hip_check(hipMemcpyAsync(system_matrix, mat_d, n * n * 36 * sizeof(double), hipMemcpyDefault, s->stream),
                      "copy");
        hip_check(hipStreamSynchronize(s->stream), "sync");
        return KYN_OK;
    });
}

}  // extern "C"
