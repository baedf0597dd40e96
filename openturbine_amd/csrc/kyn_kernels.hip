// HIP kernels of the beam Newton path for gfx950 (MI355X).
//
// Two execution forms share every device routine:
//   plain  — the reference's launch structure (src/step/step.hpp:54-83): predict, then per
//            Newton iteration tangent -> element -> constraints -> assemble -> LU/solve ->
//            update, with the convergence error read by the host after every iteration.
//   fused  — one workgroup per model runs the whole time step (predict, Newton loop,
//            epilogue) in one launch; the KKT block lives in LDS, the element tangent goes
//            from registers straight into it, nothing but state round-trips through HBM.
//
// Element contraction (the reference's integrate_{stiffness,inertia}_matrix.hpp and
// calculate_system_matrix.hpp): S_ij = sum_q phi_i R0_q[j] + phi'_i R1_q[j] where
//   R0_q[j] = (phi_j wJ(Kuu+Quu) + phi'_j w Puu) T_j + phi_j wJ(beta' Muu + gamma' Guu)
//   R1_q[j] = (phi_j w Ouu + phi'_j (w/J) Cuu) T_j
// i.e. an (N x 2Q) x (2Q x 36N) product.  FP64 tangent: each of 252 threads owns one 6x6
// component c and the nodes j = jg, jg+7, jg+14, forms its own R entries per QP in registers
// and accumulates all rows i (a 21 x 3 register tile); no LDS staging of R, no barrier inside
// the QP loop.  FP32 tangent: v_mfma_f32_16x16x4f32 tiles of the same product, every lane
// forming the one R entry of its B operand (contract_mfma_f32).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "kyn_device.hpp"
#include "kyn_system.hpp"

namespace kyn {

// Optional per-phase timestamps for profiling builds; expands to nothing in the library.
#ifndef KYN_PHASE_MARK
#define KYN_PHASE_MARK(i)
#endif
// Optional spans inside a phase (profiling builds): KYN_SPAN_BEGIN(v) takes a timestamp,
// KYN_SPAN_END(slot, v, who) adds the elapsed cycles to accumulator `slot` from thread `who`.
#ifndef KYN_SPAN_BEGIN
#define KYN_SPAN_BEGIN(v)
#define KYN_SPAN_END(slot, v, who)
#endif

constexpr int kLuTile = 144;  // row bookkeeping of the block LU (n <= kMaxBlockDofs < kLuTile)
static_assert(kMaxBlockDofs < kLuTile && kMaxBlockDofs < 192, "block LU: 3 rows per panel lane, 8-bit row keys");
constexpr int kContractThreads = 252;  // 36 components x 7 node groups
// {phi_i(q), phi'_i(q)} pairs per QP, rows padded with zeros to 22: one QP's pairs are whole
// 16-byte vectors (11 for FP32, 22 for FP64)
constexpr int kPhiStride = 44;
typedef float f32x4 __attribute__((ext_vector_type(4)));
static_assert(kPhiStride >= 2 * kMaxElemNodes && kPhiStride % 4 == 0, "phi pair stride");

// ---------------------------------------------------------------------------------------
// LDS layout (identical on host and device)
// ---------------------------------------------------------------------------------------
struct Lay {
    // persistent
    int loc, rhs, xs, xs2, piv, red, cres, misc, phys_of, lcol, lu_sync;
    // overlay: element view
    int nq, nv, nvd, tr, phi, dphi, w, J, qpF, qpM, phiT;
    // overlay: matrix view
    int A;
    int lda;
    int total;
};

__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

// Empty asm use: the value must exist here (keeps a load unconditional instead of being sunk
// into a branch by the select that follows it).  Placed after a batch of loads, it costs one
// wait for the batch.
template <typename TT>
__device__ __forceinline__ void pin(TT& v) {
    asm volatile("" : "+v"(v));
}

// cm: most constraints in one block; refine: FP64 refinement scratch needed.
__host__ __device__ inline Lay make_layout(int Nm, int Qm, int nm, int cm, bool refine, int tt_size, bool with_matrix) {
    Lay L{};
    int o = 0;
    auto take = [&](int bytes) {
        const int r = o;
        o = align16(o + bytes);
        return r;
    };
    L.loc = take(Nm * 4);
    L.rhs = take(nm * 8);
    L.xs = take(nm * 8);
    L.xs2 = take(refine ? nm * 8 : 0);
    L.piv = take(nm * 4);
    L.red = take(24 * 8);  // block reductions; in the constraint phase two 3 x 3 tangent blocks
    L.cres = take(cm * 24 * 8);  // per constraint: res, lres_b, lres_t, sys_res
    L.misc = take(16 * 4);
    L.phys_of = take(2 * kLuTile * 4);
    L.lcol = take(with_matrix ? 2 * 256 * tt_size : 0);  // two panels' L11, column-major 16 x 16
    L.lu_sync = take(16);                                  // look-ahead counter of the blocked LU
    const int overlay = o;
    L.nq = take(Nm * 7 * 8);
    L.nv = take(Nm * 6 * 8);
    L.nvd = take(Nm * 6 * 8);
    L.tr = take(Nm * 9 * 8);
    L.phi = take(Nm * Qm * 8);
    L.dphi = take(Nm * Qm * 8);
    L.w = take(Qm * 8);
    L.J = take(Qm * 8);
    L.qpF = take(Qm * 24 * 8);
    L.qpM = take(Qm * 6 * 36 * tt_size);
    L.phiT = take(Qm * kPhiStride * tt_size);  // {phi_i(q), phi'_i(q)} pairs, QP-major, tangent precision
    const int elem_end = o;
    L.lda = (nm + 1) | 1;  // [A | b]; odd row stride: row-per-lane accesses hit distinct banks
    L.A = overlay;
    const int mat_end = with_matrix ? align16(overlay + nm * L.lda * tt_size) : overlay;
    L.total = elem_end > mat_end ? elem_end : mat_end;
    return L;
}

// ---------------------------------------------------------------------------------------
// Small block-wide helpers
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.;
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; ++k) t += red[k];
    __syncthreads();
    return t;
}

template <typename T>
__device__ __forceinline__ T tabs(T v) { return v < T(0) ? -v : v; }

// ---------------------------------------------------------------------------------------
// Node-pair contraction of the FP32 tangent on the matrix cores.
//
// S_ij[r][s] = sum_q phi_i(q) R0_q[j][r][s] + phi'_i(q) R1_q[j][r][s] is the product of
// L (N x 2Q: phi_i(q), phi'_i(q)) and R (2Q x 36N), formed as v_mfma_f32_16x16x4f32 tiles: rows
// i in one or two 16-row tiles, k = 4 ks + lk covering QPs 2 ks and 2 ks + 1 (lane group lk: QP
// 2 ks + lk / 2, R0 for even lk, R1 for odd), 16 columns per tile.  Every lane forms the one R
// entry its B operand needs, from its QP's matrices (LDS) and node j's tangent block.  Columns
// are ordered c = h EP + e with h = 0 for the translational columns s < 3 (T_j is the identity
// there) and h = 1 for the rotational ones (R carries T_j), e = 18 j + 3 r + s' (s = 3 h + s'),
// EP = 16 TH, TH = ceil(18 N / 16) tiles per half.
//
// Wave w owns tiles w, w + 4, ... (tile ids are wave-uniform scalars, so the branches on them
// are scalar branches) and runs the QP loop inside each tile: a lane's column (j, r, s) and node
// j's tangent column are fixed for the whole loop, which per QP pair loads the lane's A entry,
// node j's shape pair and the QP-matrix entries of its R, forms R with 2 (h = 0) or 8 (h = 1)
// FMAs and issues one or two MFMAs.  The accumulators stay in registers until every wave is
// done with the element view, which the matrix view overlays.
// ---------------------------------------------------------------------------------------
constexpr int kMfmaTilesPerWave = 12;  // 2 halves x ceil(18 x 21 / 16) tiles over 4 waves
static_assert(2 * ((18 * kMaxElemNodes + 15) / 16) <= 4 * kMfmaTilesPerWave, "tiles per wave");

// R entry of a translational column (T_j = I): R0 = phi_j (Mi + Kuu + Quu)[r][s] + phi'_j
// Puu[r][s], R1 = phi_j Ouu[r][s] + phi'_j Cuu[r][s].  M6 points at row r of the QP's matrices
// (Kuu, Quu, Mi, Puu, Ouu, Cuu at offsets 0, 36, 72, 108, 144, 180).  Every lane loads the same
// set of entries and selects (no divergent branch around the loads).
__device__ __forceinline__ float r_entry_trans(const float* M6, bool t, int sp, float pj, float dj) {
    const float a = M6[(t ? 144 : 72) + sp];
    const float y = M6[(t ? 180 : 108) + sp];
    const float k = M6[sp] + M6[36 + sp];
    const float x = t ? a : a + k;
    return __builtin_fmaf(pj, x, dj * y);
}

// R entry of a rotational column s = 3 + s': R0 = phi_j (Mi[r][s] + sum_u (Kuu + Quu)[r][3+u]
// T[u][s']) + phi'_j sum_u Puu[r][3+u] T[u][s'], R1 = sum_u (phi_j Ouu + phi'_j Cuu)[r][3+u]
// T[u][s'], with (t0, t1, t2) = column s' of T_j.
__device__ __forceinline__ float r_entry_rot(const float* M6, bool t, int sp, float pj, float dj, float t0, float t1,
                                             float t2) {
    const float* X = M6 + (t ? 147 : 3);
    const float* Y = M6 + (t ? 183 : 111);
    const float a0 = X[0], a1 = X[1], a2 = X[2];
    const float q0 = M6[39], q1 = M6[40], q2 = M6[41];
    const float y0 = Y[0], y1 = Y[1], y2 = Y[2];
    const float m = M6[75 + sp];
    const float x0 = t ? a0 : a0 + q0, x1 = t ? a1 : a1 + q1, x2 = t ? a2 : a2 + q2;
    const float z = t ? 0.f : m;
    const float za = __builtin_fmaf(x2, t2, __builtin_fmaf(x1, t1, __builtin_fmaf(x0, t0, z)));
    const float zb = __builtin_fmaf(y2, t2, __builtin_fmaf(y1, t1, y0 * t0));
    return __builtin_fmaf(pj, za, dj * zb);
}

// Column ecol of one half of the R operand -> node j, row r and column s' of the 6 x 6 block;
// padded columns (ecol >= 18 N) map to the last column (their products are never written out).
__device__ __forceinline__ void mfma_column(int ecol, int N, int& j, int& r, int& sp) {
    const int ec = ecol < 18 * N ? ecol : 18 * N - 1;
    j = ec / 18;
    const int rem = ec - 18 * j;
    r = rem / 3;
    sp = rem - 3 * r;
}

template <bool kFused, bool kTwoRows>
__device__ __forceinline__ void contract_mfma_f32(const DevSys& S, char* sm, const Lay& L, int n, int N, int Q,
                                                  const int* loc, double* mat_out) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int Nu = __builtin_amdgcn_readfirstlane(N), Qu = __builtin_amdgcn_readfirstlane(Q);
    const int li = lane & 15, lk = lane >> 4, qo = lk >> 1;
    const bool t = (lk & 1) != 0;
    const float* qpM = reinterpret_cast<const float*>(sm + L.qpM);
    const float* phiT = reinterpret_cast<const float*>(sm + L.phiT);
    const double* tr = reinterpret_cast<const double*>(sm + L.tr);
    const int TH = (18 * Nu + 15) >> 4;
    const int npairs = Qu >> 1;  // QP pairs with both QPs present; an odd Q adds one with QP Q - 1 only
    // A operand: row li (and 16 + li) of L at column k: phi (t = 0) or phi' (t = 1) of the lane's QP
    const int ia = 2 * li + (t ? 1 : 0);
    const bool row2 = li < kPhiStride / 2 - 16;  // rows 16 + li present in the zero-padded phi table
    const int ib = row2 ? ia + 32 : ia;
    // the lane's QP within a pair, and the odd-Q tail (QP Q - 1 for both lane halves, A zero for qo = 1)
    const float* M6q = qpM + qo * 216;
    const float* phq = phiT + qo * kPhiStride;
    const int qt = Qu - 1;
    const bool tail = (Qu & 1) != 0;
    const bool tail_valid = qo == 0;
    // accumulators: one 16-row tile per column tile (N <= 16), two for N > 16
    constexpr int kNB = kTwoRows ? 2 : 1;
    f32x4 acc[kMfmaTilesPerWave][kNB];
#pragma unroll
    for (int s = 0; s < kMfmaTilesPerWave; ++s)
#pragma unroll
        for (int b = 0; b < kNB; ++b) acc[s][b] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Tiles in pairs within one half: tile pair (s', s' + 1) of half h runs one QP loop for
    // both tiles -- two independent MFMA chains per QP pair, the loads of one tile overlapping
    // the other's arithmetic.  acc[h * kTilesPerHalf + s'] holds tile wv + 4 s' of half h.
    constexpr int kTilesPerHalf = kMfmaTilesPerWave / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int s2 = 0; s2 < kTilesPerHalf; s2 += 2) {
            const int TA = wv + 4 * s2;  // the pair's second tile, TA + 4, may lie past the half:
            if (TA < TH) {               // it then forms clamped columns that are never written out
                const int sA = h * kTilesPerHalf + s2;
                int jA, rA, pA, jB, rB, pB;
                mfma_column(16 * TA + li, Nu, jA, rA, pA);
                mfma_column(16 * (TA + 4) + li, Nu, jB, rB, pB);
                f32x4 a0 = acc[sA][0], a1 = acc[sA][kNB - 1], b0 = acc[sA + 1][0], b1 = acc[sA + 1][kNB - 1];
                if (h == 0) {
                    for (int ks = 0; ks < npairs; ++ks) {
                        const float* ph = phq + ks * (2 * kPhiStride);
                        const float* M6 = M6q + ks * 432;
                        const float RA = r_entry_trans(M6 + rA * 6, t, pA, ph[2 * jA], ph[2 * jA + 1]);
                        const float RB = r_entry_trans(M6 + rB * 6, t, pB, ph[2 * jB], ph[2 * jB + 1]);
                        const float x0 = ph[ia], y1 = ph[ib];
                        const float x1 = row2 ? y1 : 0.f;
                        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RA, a0, 0, 0, 0);
                        b0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RB, b0, 0, 0, 0);
                        if constexpr (kTwoRows) {
                            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RA, a1, 0, 0, 0);
                            b1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RB, b1, 0, 0, 0);
                        }
                    }
                    if (tail) {
                        const float* ph = phiT + qt * kPhiStride;
                        const float* M6 = qpM + qt * 216;
                        const float RA = r_entry_trans(M6 + rA * 6, t, pA, ph[2 * jA], ph[2 * jA + 1]);
                        const float RB = r_entry_trans(M6 + rB * 6, t, pB, ph[2 * jB], ph[2 * jB + 1]);
                        const float y0 = ph[ia], y1 = ph[ib];
                        const float x0 = tail_valid ? y0 : 0.f, x1 = tail_valid && row2 ? y1 : 0.f;
                        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RA, a0, 0, 0, 0);
                        b0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RB, b0, 0, 0, 0);
                        if constexpr (kTwoRows) {
                            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RA, a1, 0, 0, 0);
                            b1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RB, b1, 0, 0, 0);
                        }
                    }
                } else {
                    // column s' of node j's rotational tangent block, per tile
                    const double* TjA = tr + jA * 9 + pA;
                    const double* TjB = tr + jB * 9 + pB;
                    const float tA0 = (float)TjA[0], tA1 = (float)TjA[3], tA2 = (float)TjA[6];
                    const float tB0 = (float)TjB[0], tB1 = (float)TjB[3], tB2 = (float)TjB[6];
                    for (int ks = 0; ks < npairs; ++ks) {
                        const float* ph = phq + ks * (2 * kPhiStride);
                        const float* M6 = M6q + ks * 432;
                        const float RA =
                            r_entry_rot(M6 + rA * 6, t, pA, ph[2 * jA], ph[2 * jA + 1], tA0, tA1, tA2);
                        const float RB =
                            r_entry_rot(M6 + rB * 6, t, pB, ph[2 * jB], ph[2 * jB + 1], tB0, tB1, tB2);
                        const float x0 = ph[ia], y1 = ph[ib];
                        const float x1 = row2 ? y1 : 0.f;
                        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RA, a0, 0, 0, 0);
                        b0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RB, b0, 0, 0, 0);
                        if constexpr (kTwoRows) {
                            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RA, a1, 0, 0, 0);
                            b1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RB, b1, 0, 0, 0);
                        }
                    }
                    if (tail) {
                        const float* ph = phiT + qt * kPhiStride;
                        const float* M6 = qpM + qt * 216;
                        const float RA =
                            r_entry_rot(M6 + rA * 6, t, pA, ph[2 * jA], ph[2 * jA + 1], tA0, tA1, tA2);
                        const float RB =
                            r_entry_rot(M6 + rB * 6, t, pB, ph[2 * jB], ph[2 * jB + 1], tB0, tB1, tB2);
                        const float y0 = ph[ia], y1 = ph[ib];
                        const float x0 = tail_valid ? y0 : 0.f, x1 = tail_valid && row2 ? y1 : 0.f;
                        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RA, a0, 0, 0, 0);
                        b0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, RB, b0, 0, 0, 0);
                        if constexpr (kTwoRows) {
                            a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RA, a1, 0, 0, 0);
                            b1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, RB, b1, 0, 0, 0);
                        }
                    }
                }
                acc[sA][0] = a0;
                acc[sA + 1][0] = b0;
                if constexpr (kTwoRows) {
                    acc[sA][kNB - 1] = a1;
                    acc[sA + 1][kNB - 1] = b1;
                }
            }
        }
    }
    __syncthreads();  // every wave done with the element view
    KYN_PHASE_MARK(4);
    // ---- write out: lane holds S[i = 16 b + 4 lk + r][column of its tile]
    float* A = reinterpret_cast<float*>(sm + L.A);
    const float cond = (float)S.p.conditioner;
    int rowoff[2][4];  // block-local row of node i (fused) or i (otherwise); -1 past N
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * b + 4 * lk + r;
            rowoff[b][r] = i < Nu ? (kFused ? loc[i] : i) : -1;
        }
    if (kFused) {
        uint4* A4 = reinterpret_cast<uint4*>(sm + L.A);
        const int nvec = (n * L.lda * (int)sizeof(float) + 15) / 16;
        for (int k = tid; k < nvec; k += blockDim.x) A4[k] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kMfmaTilesPerWave; ++s) {
        const int h = s / kTilesPerHalf;
        const int T = wv + 4 * (s % kTilesPerHalf);  // tile within half h
        if (T < TH) {
            const int ecol = 16 * T + li;
            if (ecol < 18 * Nu) {
                const int j = ecol / 18, rem = ecol - 18 * j, r_ = rem / 3, col = 3 * h + rem - 3 * r_;
                if (kFused) {
                    const int c = loc[j] + col;
#pragma unroll
                    for (int b = 0; b < kNB; ++b)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (rowoff[b][r] >= 0) A[(rowoff[b][r] + r_) * L.lda + c] = acc[s][b][r] * cond;
                } else {
                    double* out = mat_out + (size_t)j * 36 + r_ * 6 + col;
#pragma unroll
                    for (int b = 0; b < kNB; ++b)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (rowoff[b][r] >= 0) out[(size_t)rowoff[b][r] * Nu * 36] = (double)acc[s][b][r];
                }
            }
        }
    }
    if (kFused) __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Node-pair contraction on the VALU (the FP64 tangent's, and the FP32 tangent's where the tuning
// table picks it for the element size).
// ---------------------------------------------------------------------------------------
template <typename TT, bool kFused>
__device__ __forceinline__ void contract_valu(const DevSys& S, char* sm, const Lay& L, int n, int N, int Q,
                                              const int* loc, double* mat_out) {
    const int tid = threadIdx.x;
    const TT* qpM = reinterpret_cast<const TT*>(sm + L.qpM);
    const TT* phiT = reinterpret_cast<const TT*>(sm + L.phiT);
    const double* tr = reinterpret_cast<const double*>(sm + L.tr);
    // ---- contraction, one pass over the QPs, no barrier
    // Thread t < 252 owns component cc = t % 36 = (r, s) and nodes j = t / 36 + 7 b (b < 3);
    // per QP it forms its own R0_q[j][cc], R1_q[j][cc] from row r of the QP matrices and
    // column s of T_j, then accumulates S_ij[cc] += phi_i R0 + phi'_i R1 for every row i
    // (the QP's phi pairs: broadcast vector loads, all rows at once).  Each R entry is formed by
    // exactly one thread.
    const int cc = tid % 36, jg = tid / 36;
    const bool active = tid < kContractThreads;
    const int r_ = cc / 6, s_ = cc - 6 * (cc / 6);
    const bool rot = s_ >= 3;
    const int col0 = rot ? 3 : 0;  // columns of A/P/O/C combined with T_j's column (s - 3)
    TT tc[3][3];
    int jc[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int j = jg + 7 * b;
        const bool ok = active && j < N;
        jc[b] = ok ? j : 0;
#pragma unroll
        for (int u = 0; u < 3; ++u)
            tc[b][u] = !ok ? TT(0) : (rot ? (TT)tr[j * 9 + u * 3 + (s_ - 3)] : (u == s_ ? TT(1) : TT(0)));
    }
    TT acc[kMaxElemNodes][3];
#pragma unroll
    for (int i = 0; i < kMaxElemNodes; ++i)
#pragma unroll
        for (int b = 0; b < 3; ++b) acc[i][b] = TT(0);
    const int Nu = __builtin_amdgcn_readfirstlane(N);  // uniform: a scalar branch per row group
    for (int q = 0; q < Q; ++q) {
        const TT* M6 = qpM + (size_t)q * 6 * 36 + r_ * 6;
        TT ra[3], rp[3], ro[3], rc[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            ra[u] = M6[col0 + u] + M6[36 + col0 + u];
            rp[u] = M6[108 + col0 + u];
            ro[u] = M6[144 + col0 + u];
            rc[u] = M6[180 + col0 + u];
        }
        const TT rm = M6[72 + s_];
        const TT* ph = phiT + (size_t)q * kPhiStride;
        TT pj[3], dj[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            pj[b] = ph[jc[b] * 2];
            dj[b] = ph[jc[b] * 2 + 1];
        }
        // every row's pair for this QP in registers (whole vectors, one wait), then the rows
        constexpr int VW = 16 / (int)sizeof(TT);
        typedef TT vec_t __attribute__((ext_vector_type(16 / sizeof(TT))));
        TT phv[kPhiStride];
#pragma unroll
        for (int c = 0; c < kPhiStride / VW; ++c) {
            const vec_t v = *reinterpret_cast<const vec_t*>(ph + c * VW);
#pragma unroll
            for (int e = 0; e < VW; ++e) phv[c * VW + e] = v[e];
        }
        TT R0[3], R1[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            TT v0 = pj[b] * rm, v1 = TT(0);
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                v0 += (pj[b] * ra[u] + dj[b] * rp[u]) * tc[b][u];
                v1 += (pj[b] * ro[u] + dj[b] * rc[u]) * tc[b][u];
            }
            R0[b] = v0;
            R1[b] = v1;
        }
        // rows in groups of three behind a scalar branch on N (rows past N are skipped)
#pragma unroll
        for (int i0 = 0; i0 < kMaxElemNodes; i0 += 3) {
            if (i0 < Nu) {
#pragma unroll
                for (int i = i0; i < i0 + 3 && i < kMaxElemNodes; ++i) {
                    const TT pi = phv[i * 2], di = phv[i * 2 + 1];
#pragma unroll
                    for (int b = 0; b < 3; ++b) acc[i][b] += pi * R0[b] + di * R1[b];
                }
            }
        }
    }
    __syncthreads();

    KYN_PHASE_MARK(4);
    // ---- write out
    if (kFused) {
        // the overlay is free after the barrier above: zero the matrix view, then write
        TT* A = reinterpret_cast<TT*>(sm + L.A);
        {   // 16-byte stores; the matrix view is 16-byte aligned and reserved to a 16-byte multiple
            uint4* A4 = reinterpret_cast<uint4*>(sm + L.A);
            const int nv = (n * L.lda * (int)sizeof(TT) + 15) / 16;
            for (int k = tid; k < nv; k += blockDim.x) A4[k] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        const TT cond = (TT)S.p.conditioner;
        if (active) {
#pragma unroll
            for (int i = 0; i < kMaxElemNodes; ++i) {
                if (i < N) {
                    const int row = loc[i] + r_;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const int j = jg + 7 * b;
                        if (j < N) A[row * L.lda + loc[j] + s_] = acc[i][b] * cond;
                    }
                }
            }
        }
        __syncthreads();
    } else {
        if (active) {
            double* out = mat_out;
#pragma unroll
            for (int i = 0; i < kMaxElemNodes; ++i) {
                if (i < N) {
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const int j = jg + 7 * b;
                        if (j < N) out[(size_t)(i * N + j) * 36 + cc] = (double)acc[i][b];
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Element phase: gather, QP values, residual, contraction.
//   kFused: residual goes into the block rhs (-f + R) and the tangent into LDS matrix A
//           (times the conditioner, at block-local dofs); otherwise into res_out ([N][6]) and
//           mat_out ([N][N][36]).
//   kOwnTangent: the nodes' rotational tangent blocks are formed here from q_delta
//           (UpdateTangentOperator folded in) rather than read from tan_rot.
// ---------------------------------------------------------------------------------------
//   model: the element's model where the caller knows it (one load for the gravity vector
//           instead of the element -> node -> block -> model chain), else -1.
template <typename TT, bool kFused, bool kOwnTangent = kFused>
__device__ __forceinline__ void element_phase(const DevSys& S, int e, char* sm, const Lay& L, int n,
                                              double* res_out = nullptr, double* mat_out = nullptr,
                                              int model = -1) {
    const int tid = threadIdx.x;
    const int N = S.elem_N[e], Q = S.elem_Q[e];
    const int noff = S.elem_node_off[e];
    const int qoff = S.elem_qp_off[e];
    const int soff = S.elem_shape_off[e];
    const double* grav =
        S.model_gravity + 3 * (model >= 0 ? model : S.block_model[S.node_block[S.elem_nodes[noff]]]);
    double* nq = reinterpret_cast<double*>(sm + L.nq);
    double* nv = reinterpret_cast<double*>(sm + L.nv);
    double* nvd = reinterpret_cast<double*>(sm + L.nvd);
    double* tr = reinterpret_cast<double*>(sm + L.tr);
    int* loc = reinterpret_cast<int*>(sm + L.loc);
    double* phi = reinterpret_cast<double*>(sm + L.phi);
    double* dphi = reinterpret_cast<double*>(sm + L.dphi);
    double* wq = reinterpret_cast<double*>(sm + L.w);
    double* Jq = reinterpret_cast<double*>(sm + L.J);
    double* qpF = reinterpret_cast<double*>(sm + L.qpF);
    TT* qpM = reinterpret_cast<TT*>(sm + L.qpM);
    TT* phiT = reinterpret_cast<TT*>(sm + L.phiT);
    double* rhs = reinterpret_cast<double*>(sm + L.rhs);
    const double h = S.p.h;

    // ---- gather node state (update_node_state.hpp:26-35) and tangent blocks on the last wave,
    // the shape and QP tables on the waves before it: the two load chains run side by side
    const int nodeT0 = blockDim.x > 64 ? (int)blockDim.x - 64 : 0;  // first thread of the node gather
    const int tabT = blockDim.x > 64 ? nodeT0 : (int)blockDim.x;     // threads of the table gather
    for (int k = tid - nodeT0; k >= 0 && k < N; k += blockDim.x - nodeT0) {
        const int node = S.elem_nodes[noff + k];
#pragma unroll
        for (int c = 0; c < 7; ++c) nq[k * 7 + c] = S.q[node * 7 + c];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            nv[k * 6 + c] = S.v[node * 6 + c];
            nvd[k * 6 + c] = S.vd[node * 6 + c];
        }
        if (kOwnTangent) {
            // UpdateTangentOperator (update_tangent_operator.hpp:13-26) folded in
            double qd[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) qd[c] = S.q_delta[node * 6 + c];
            tangent_rot_block(h, qd, tr + k * 9);
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) tr[k * 9 + c] = S.tan_rot[node * 9 + c];
        }
        loc[k] = kFused ? S.node_local[node] : 0;
    }
    for (int k = tid; tid < tabT && k < (kPhiStride / 2) * Q; k += tabT) {
        const int i = k / Q, q = k - Q * (k / Q);
        double p = 0., dp = 0.;
        if (i < N) {
            p = S.shape_interp[soff + k];
            dp = S.shape_deriv[soff + k];
            phi[k] = p;
            dphi[k] = dp;
        }
        phiT[q * kPhiStride + 2 * i] = (TT)p;
        phiT[q * kPhiStride + 2 * i + 1] = (TT)dp;
    }
    for (int q = tid; tid < tabT && q < Q; q += tabT) {
        wq[q] = S.qp_weight[qoff + q];
        Jq[q] = S.qp_jac[qoff + q];
    }
    __syncthreads();
    KYN_PHASE_MARK(1);

    // ---- QP values: wave 0 the inertial terms, wave 1 the stiffness terms, lane <-> QP (Q <= 64:
    // each wave runs one of the two code paths once, and the two waves run side by side)
    static_assert(kMaxElemQps <= 64, "one lane per QP");
    if (tid < 128 && (tid & 63) < Q) {
        const int q = tid & 63;
        const int half = tid >> 6;
        const double w = wq[q], J = Jq[q];
        const double* r0 = S.qp_r0 + (size_t)(qoff + q) * 4;
        TT* M6 = qpM + (size_t)q * 6 * 36;  // A1, A2, Mi, P, O, C
        double* F = qpF + q * 24;          // Fc, Fd, Fi, Fg
        if (half == 0) {
            // interpolate_to_quadrature_point_for_inertia.hpp:19-40
            double r[4] = {0., 0., 0., 0.}, ud[3] = {0., 0., 0.}, om[3] = {0., 0., 0.},
                   omd[3] = {0., 0., 0.};
            for (int k = 0; k < N; ++k) {
                const double p = phi[k * Q + q];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ud[c] += nvd[k * 6 + c] * p;
                    om[c] += nv[k * 6 + 3 + c] * p;
                    omd[c] += nvd[k * 6 + 3 + c] * p;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) r[c] += nq[k * 7 + 3 + c] * p;
            }
            double r0l[4] = {r0[0], r0[1], r0[2], r0[3]};
            qp_inertial<TT>(r, ud, om, omd, r0l, S.qp_Mstar + (size_t)(qoff + q) * 36, grav, w * J, S.p.beta_prime,
                            S.p.gamma_prime, M6, F);
        } else {
            // interpolate_to_quadrature_point_for_stiffness.hpp:19-41
            double r[4] = {0., 0., 0., 0.}, rp[4] = {0., 0., 0., 0.}, up[3] = {0., 0., 0.};
            for (int k = 0; k < N; ++k) {
                const double p = phi[k * Q + q];
                const double dp = dphi[k * Q + q] / J;
#pragma unroll
                for (int c = 0; c < 3; ++c) up[c] += nq[k * 7 + c] * dp;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    r[c] += nq[k * 7 + 3 + c] * p;
                    rp[c] += nq[k * 7 + 3 + c] * dp;
                }
            }
            const double r0l[4] = {r0[0], r0[1], r0[2], r0[3]};
            const double* xp = S.qp_x0p + (size_t)(qoff + q) * 3;
            const double x0p[3] = {xp[0], xp[1], xp[2]};
            qp_stiffness<TT>(r, rp, up, r0l, x0p, S.qp_Cstar + (size_t)(qoff + q) * 36, w * J, w, w / J, M6, F);
        }
    }
    __syncthreads();
    KYN_PHASE_MARK(2);

    // ---- residual (integrate_residual_vector.hpp:32-50)
    for (int t = tid; t < 6 * N; t += blockDim.x) {
        const int k = t / 6, c = t - 6 * (t / 6);
        // HBM operands first: their latency runs under the QP loop
        const double fx = S.node_FX[(size_t)(noff + k) * 6 + c];
        const double fk = kFused ? S.f[S.elem_nodes[noff + k] * 6 + c] : 0.;
        double acc = 0.;
        for (int q = 0; q < Q; ++q) {
            const double w = wq[q];
            const double cc = w * dphi[k * Q + q];
            const double cd = w * Jq[q] * phi[k * Q + q];
            const double* F = qpF + q * 24;
            acc += cc * F[c] + cd * ((F[6 + c] + F[12 + c]) - F[18 + c]);
        }
        const double R = acc - fx;
        if (kFused) {
            rhs[loc[k] + c] = -fk + R;
        } else {
            res_out[k * 6 + c] = R;
        }
    }

    KYN_PHASE_MARK(3);
    if constexpr (sizeof(TT) == 4) {
        // FP32 tangent: the matrix cores or the VALU register tile, per element size (the
        // tuning table: bit N of contract_mfma_mask)
        if ((S.contract_mfma_mask >> N) & 1u) {
            if (N > 16)
                contract_mfma_f32<kFused, true>(S, sm, L, n, N, Q, loc, mat_out);
            else
                contract_mfma_f32<kFused, false>(S, sm, L, n, N, Q, loc, mat_out);
        } else
            contract_valu<TT, kFused>(S, sm, L, n, N, Q, loc, mat_out);
    } else {
        contract_valu<TT, kFused>(S, sm, L, n, N, Q, loc, mat_out);
    }
}

// ---------------------------------------------------------------------------------------
// Blocked LU with partial pivoting on an augmented LDS matrix [A | b] (replaces
// dss::numeric_factorization / dss::solve, src/step/solve_system.hpp:27-31).
//
// Right-looking, panels of 16 columns.  Wave 0 factors a panel in registers (lane l holds the
// rows not yet pivoted, logical c0 + l, c0 + l + 64, ...): per column a candidate per lane, a
// 64-lane argmax by DPP row exchanges and row broadcasts (FP64: max, then the lowest logical row
// holding it; FP32: one reduction of a packed magnitude / row key), the pivot row broadcast by
// readlane, then the in-panel rank-1 update; no barrier inside a panel.  The other waves form
// the panel's U rows over the trailing columns (forward substitution with the panel's unit L)
// and update the trailing rows, the next panel's columns first (look-ahead).  Every entry sees
// the same FMA sequence as in the oracle's unblocked loop (a_ij -= l_ik u_kj for k ascending),
// so FP64 factors are bit-identical to the oracle's.
//
// Pivoting is implicit: rows never move; phys_of maps logical -> physical rows (the position
// the oracle's dense LU holds a row at after its swaps).  Pivot: largest |a_ik| over
// logical rows i >= k, lowest logical index on ties, row k kept when a_kk is NaN.  The right-
// hand side b is column n, so elimination applies P and L^-1 to it on the way.
// On return: L (unit) and U at physical rows, y = L^-1 P b in column n, piv as the oracle's.
// ---------------------------------------------------------------------------------------
constexpr int kPanel = 16;

struct LuScratch {
    int* phys_of;  // physical row of each logical row: two buffers of kLuTile, the mapping after
                   // panel p in buffer p & 1 (the next panel's factorization writes the other
                   // buffer while this panel is still being applied); after the LU, buffer 0
    void* lcol;    // two 16 x 16 column-major L11 tiles (panel p in tile p & 1), storage type
    int* sync;     // look-ahead counter: waves 1.. add one each when the next panel's columns are done
    int wait_limit;  // polls before wave 0 gives up waiting on the counter (< 0: at once)
    int* error;      // device error flags (bit kDevErrLuWait set on giving up)
};
__device__ __forceinline__ int* phys_buf(const LuScratch& W, int c0) {
    return W.phys_of + ((c0 / kPanel) & 1) * kLuTile;
}
template <typename TT>
__device__ __forceinline__ TT* panel_l11(const LuScratch& W, int c0) {
    return reinterpret_cast<TT*>(W.lcol) + ((c0 / kPanel) & 1) * kPanel * kPanel;
}

constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppQuad3210 = 0x1B, kDppQuad1032 = 0xB1;

template <int C>
__device__ __forceinline__ int dpp_x(int v) {
    return __builtin_amdgcn_update_dpp(0, v, C, 0xF, 0xF, false);
}
template <int C>
__device__ __forceinline__ int lane_x(int v) {
    return dpp_x<C>(v);
}
template <int C>
__device__ __forceinline__ float lane_x(float v) {
    return __int_as_float(lane_x<C>(__float_as_int(v)));
}
template <int C>
__device__ __forceinline__ double lane_x(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_x<C>((int)(b & 0xffffffffll)), hi = lane_x<C>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Row broadcasts (GFX9 DPP): lane 15 of each row to the next row (written rows 1 and 3), and
// lane 31 to rows 2 and 3; rows outside the mask keep `old`.
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
template <int C, int RM>
__device__ __forceinline__ int dpp_old(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, C, RM, 0xF, false);
}
template <int C, int RM>
__device__ __forceinline__ float dpp_old(float old, float v) {
    return __int_as_float(dpp_old<C, RM>(__float_as_int(old), __float_as_int(v)));
}
template <int C, int RM>
__device__ __forceinline__ double dpp_old(double old, double v) {
    const long long o = __double_as_longlong(old), b = __double_as_longlong(v);
    const int lo = dpp_old<C, RM>((int)(o & 0xffffffffll), (int)(b & 0xffffffffll));
    const int hi = dpp_old<C, RM>((int)(o >> 32), (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Reduction over the 64 lanes whose result is valid in lane 63: four in-row exchanges, then the
// two row broadcasts (no cross-row swaps).
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce_to63(T v, Op op) {
    v = op(v, lane_x<kDppQuad1032>(v));
    v = op(v, lane_x<kDppQuad3210>(v));
    v = op(v, lane_x<kDppRowHalfMirror>(v));
    v = op(v, lane_x<kDppRowMirror>(v));
    v = op(v, dpp_old<kDppRowBcast15, 0xA>(v, v));
    return op(v, dpp_old<kDppRowBcast31, 0xC>(v, v));
}
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
// max over lanes (to every lane through lane 63)
__device__ __forceinline__ int wave_max(int v) {
    return __builtin_amdgcn_readlane(wave_reduce_to63(v, imax), 63);
}
__device__ __forceinline__ double wave_max(double v) {
    const double r = wave_reduce_to63(v, [](double a, double b) { return fmax(a, b); });
    const long long b = __double_as_longlong(r);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ int wave_min(int v) {
    return __builtin_amdgcn_readlane(wave_reduce_to63(v, imin), 63);
}
// pivot candidate magnitude: FP32 as the bit pattern of |v| (an int ordered like |v| for
// non-negative floats, so the reduction is an integer max), FP64 as |v|; -1 = not a candidate
__device__ __forceinline__ int cand_of(float v, bool ok) { return ok && v == v ? __float_as_int(fabsf(v)) : -1; }
__device__ __forceinline__ double cand_of(double v, bool ok) { return ok ? fabs(v) : -1.; }
template <typename TT>
struct Cand;
template <>
struct Cand<float> {
    using type = int;
};
template <>
struct Cand<double> {
    using type = double;
};

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Empty asm use tagged with the register slot: keeps per-slot branches from being merged into
// a dynamically indexed (scratch) register read.
template <int S, typename TT>
__device__ __forceinline__ TT slot_tag(TT v) {
    asm volatile("; lu slot %1" : "+v"(v) : "n"(S));
    return v;
}

// One pivot column (J) of the wave-0 panel factorization: pivot search over the 64 lanes, pivot
// row broadcast, multipliers and the rank-1 update of the panel's later columns.
template <typename TT, int NS, bool kFull, int J>
__device__ __forceinline__ void panel_columns(TT (&P)[NS][kPanel], int (&lg)[NS], int (&pl)[kPanel], int c0, int w,
                                              int lane) {
    if constexpr (J < kPanel) {
        if (kFull || J < w) {
            const int k = c0 + J;
            int key;
            if constexpr (sizeof(TT) == 4) {
                // FP32: one reduction of (|a_rk| to 15 mantissa bits, 255 - logical row): the
                // largest magnitude, and among magnitudes equal to 2^-15 the lowest logical row
                // (NaN never wins; no candidate or a NaN a_kk keeps row k); the physical row of
                // the winner from a ballot
                int bk = -1;
                bool dnan = false;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const TT v = P[s][J];
                    const int c = (((__float_as_int(fabsf(v)) >> 8) << 8) | (255 - lg[s]));
                    bk = (lg[s] >= k && v == v && c > bk) ? c : bk;
                    dnan |= (lg[s] == k) & (v != v);
                }
                bk = wave_max(bk);
                const int pl0 = (bk < 0 || __any(dnan)) ? k : 255 - (bk & 255);
                int kp = 0;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const unsigned long long m = __ballot(lg[s] == pl0);
                    kp = m ? (__builtin_ctzll(m) + 64 * s) : kp;
                }
                key = (pl0 << 8) | kp;
            } else {
                // largest |a_rk| over eligible rows (NaN never wins), then the lowest key holding
                // it; both reduced over the 64 lanes, so every lane sees the same pivot
                using CT = typename Cand<TT>::type;
                CT vs[NS];
                CT bv = CT(-1);
                bool dnan = false;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const TT v = P[s][J];
                    vs[s] = cand_of(v, lg[s] >= k);
                    bv = vs[s] > bv ? vs[s] : bv;
                    dnan |= (lg[s] == k) & (v != v);
                }
                bv = wave_max(bv);
                int bk = 0x7fffffff;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const int kk = (lg[s] << 8) | (lane + 64 * s);
                    bk = ((lg[s] >= k) & (vs[s] == bv)) ? (kk < bk ? kk : bk) : bk;
                }
                key = wave_min(bk);
                if (__any(dnan)) {  // the oracle keeps row k when a_kk is NaN
                    int kp = 0;
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const unsigned long long m = __ballot(lg[s] == k);
                        kp = m ? (__builtin_ctzll(m) + 64 * s) : kp;
                    }
                    key = (k << 8) | kp;
                }
            }
            const int plog = key >> 8, pphys = key & 255;
            const int ol = pphys & 63, slot = pphys >> 6;
            // pivot row (panel part) to every lane (a uniform branch on the slot: fewer
            // instructions than selecting among the slots in every lane)
            TT prow[kPanel];
#pragma unroll
            for (int jj = 0; jj < kPanel; ++jj) prow[jj] = TT(0);
            if (NS == 1 || slot == 0) {
#pragma unroll
                for (int jj = J; jj < kPanel; ++jj) prow[jj] = rdlane(P[0][jj], ol);
            } else if (NS == 2 || slot == 1) {
#pragma unroll
                for (int jj = J; jj < kPanel; ++jj) prow[jj] = rdlane(P[NS > 1 ? 1 : 0][jj], ol);
            } else {
#pragma unroll
                for (int jj = J; jj < kPanel; ++jj) prow[jj] = rdlane(P[NS > 2 ? 2 : 0][jj], ol);
            }
            const TT d = prow[J];
            pl[J] = plog;
#pragma unroll
            for (int s = 0; s < NS; ++s) lg[s] = lg[s] == k ? plog : (lg[s] == plog ? k : lg[s]);
            if (d != TT(0)) {  // a zero pivot leaves the panel unchanged (the oracle skips the column)
                const TT inv = TT(1) / d;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const bool below = lg[s] > k;
                    const TT l = below ? P[s][J] * inv : TT(0);
                    P[s][J] = below ? l : P[s][J];
#pragma unroll
                    for (int jj = J + 1; jj < kPanel; ++jj) P[s][jj] -= l * prow[jj];
                }
            }
        }
        panel_columns<TT, NS, kFull, J + 1>(P, lg, pl, c0, w, lane);
    }
}

// Factor panel columns [c0, c0 + w) in wave 0 over the rows not yet pivoted: logical rows
// c0 .. n - 1 (NS = row slots per lane, n - c0 <= 64 NS), at the physical rows of the previous
// panel's mapping.  Row bookkeeping stays in registers (lg per lane row) and is written back once
// at the end of the panel into this panel's mapping buffer, together with the panel's unit-lower
// L11 as a compact column-major tile.
template <typename TT, int NS, bool kFull>
__device__ __forceinline__ void lu_panel_ns(TT* A, int n, int lda, int c0, int w, int* piv, const LuScratch& W) {
    if (kFull) w = kPanel;
    const int lane = threadIdx.x & 63;
    const int* ph_in = phys_buf(W, c0 + kPanel);  // the previous panel's buffer (identity before panel 0)
    int* ph_out = phys_buf(W, c0);
    TT P[NS][kPanel];
    int lg[NS], rr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int li = c0 + lane + 64 * s;
        const bool ok = li < n;
        lg[s] = ok ? li : -1;
        rr[s] = ph_in[ok ? li : c0];  // clamped: loads are unconditional, values selected after
#pragma unroll
        for (int jj = 0; jj < kPanel; ++jj) P[s][jj] = A[rr[s] * lda + c0 + (jj < w ? jj : 0)];
#pragma unroll
        for (int jj = 0; jj < kPanel; ++jj) pin(P[s][jj]);  // after all loads: one wait
#pragma unroll
        for (int jj = 0; jj < kPanel; ++jj) P[s][jj] = (ok && jj < w) ? P[s][jj] : TT(0);
    }
    // Columns as compile-time steps (constant register indices without relying on unrolling:
    // the tile stays in registers).  A single wave issues about one instruction per four cycles,
    // so each step keeps its instruction count down: uniform branches on the pivot's slot and
    // on a zero pivot rather than per-element selects, and only NS row slots per lane.
    int pl[kPanel];
    panel_columns<TT, NS, kFull, 0>(P, lg, pl, c0, w, lane);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kPanel; ++j)
            if (kFull || j < w) piv[c0 + j] = pl[j];
    }
    for (int i = lane; i < c0; i += 64) ph_out[i] = ph_in[i];  // rows pivoted before this panel
    TT* Lc = panel_l11<TT>(W, c0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (lg[s] >= 0) {
#pragma unroll
            for (int jj = 0; jj < kPanel; ++jj)
                if (jj < w) A[rr[s] * lda + c0 + jj] = P[s][jj];
            ph_out[lg[s]] = rr[s];
            const int i = lg[s] - c0;  // row of L11 this lane holds, if any
            if (i < w) {
#pragma unroll
                for (int k = 0; k < kPanel; ++k) Lc[k * kPanel + i] = k < i ? P[s][k] : TT(0);
            }
        }
    }
}
template <typename TT>
__device__ __forceinline__ void lu_panel(TT* A, int n, int lda, int c0, int w, int* piv, const LuScratch& W) {
    const int rows = n - c0;
    if (w < kPanel) {
        if (rows <= 64)
            lu_panel_ns<TT, 1, false>(A, n, lda, c0, w, piv, W);
        else if (rows <= 128)
            lu_panel_ns<TT, 2, false>(A, n, lda, c0, w, piv, W);
        else
            lu_panel_ns<TT, 3, false>(A, n, lda, c0, w, piv, W);
    } else if (rows <= 64) {
        lu_panel_ns<TT, 1, true>(A, n, lda, c0, w, piv, W);
    } else if (rows <= 128) {
        lu_panel_ns<TT, 2, true>(A, n, lda, c0, w, piv, W);
    } else {
        lu_panel_ns<TT, 3, true>(A, n, lda, c0, w, piv, W);
    }
}



// Columns KG .. KG + 3 of the U-row solve: u_i -= l_ik u_k (i > k).  All 16-byte loads of the
// four L11 columns are issued first, then one wait, then the FMAs.
template <typename TT, int KG>
__device__ __forceinline__ void u_rows_group(const TT* Lc, TT* u) {
    constexpr int VW = 16 / sizeof(TT);
    typedef TT vec_t __attribute__((ext_vector_type(16 / sizeof(TT))));
    TT lk[4][kPanel];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
        for (int q = 0; q < kPanel / VW; ++q) {
            if (KG + kk < kPanel - 1 && (q + 1) * VW > KG + kk + 1) {
                const vec_t v = *reinterpret_cast<const vec_t*>(Lc + (KG + kk) * kPanel + q * VW);
#pragma unroll
                for (int e = 0; e < VW; ++e) lk[kk][q * VW + e] = v[e];
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = KG + kk + 1; i < kPanel; ++i) pin(lk[kk][i]);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = KG + kk + 1; i < kPanel; ++i) u[i] -= lk[kk][i] * u[KG + kk];
}

// The same four columns for the unpivoted FP32 LU, with explicit FMAs: u_i = fma(-l_ik, u_k,
// u_i), k ascending for every i (the order of the row-oriented loop).
template <int KG>
__device__ __forceinline__ void u_rows_group_fma(const float* Lc, float (&u)[kPanel]) {
    typedef float vec_t __attribute__((ext_vector_type(4)));
    float lk[4][kPanel];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
        for (int q = 0; q < kPanel / 4; ++q) {
            if (KG + kk < kPanel - 1 && (q + 1) * 4 > KG + kk + 1) {
                const vec_t v = *reinterpret_cast<const vec_t*>(Lc + (KG + kk) * kPanel + q * 4);
#pragma unroll
                for (int e = 0; e < 4; ++e) lk[kk][q * 4 + e] = v[e];
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = KG + kk + 1; i < kPanel; ++i) u[i] = __builtin_fmaf(-lk[kk][i], u[KG + kk], u[i]);
}

// Applies panel [c0, c0 + w) to this wave's share of the columns [ja, jb): columns
// ja + g, ja + g + gs, ... (wave g of gs).  First the panel's U rows (U12 = L11^-1 A12, lane <->
// column, written in place into the pivot rows), then the trailing rows a_rj -= sum_k l_rk u_kj
// (k ascending; lane <-> logical rows j0 + lane + 64 s, s < NSA, with their multipliers in
// registers; two columns at a time).  Rows already pivoted are not touched.
// Runs within one wave: its LDS writes are seen by its own later reads without a barrier.
template <typename TT, bool kFull, int NSA>
__device__ __forceinline__ void lu_apply_panel(TT* A, int n, int lda, int c0, int w, int ja, int jb, int g, int gs,
                                               const LuScratch& W) {
    if (kFull) w = kPanel;
    const int lane = threadIdx.x & 63;
    const int j0 = c0 + w;  // rows with logical index >= j0 are trailing rows
    const int m = ja + g < jb ? (jb - ja - g + gs - 1) / gs : 0;
    if (m == 0) return;
    // the row mapping after this panel: the next panel's factorization (wave 0, possibly
    // running meanwhile) permutes rows among logical j0 .. n - 1 only and writes the other buffer
    const int* ph = phys_buf(W, c0);
    int pr[kPanel];
#pragma unroll
    for (int i = 0; i < kPanel; ++i) pr[i] = __builtin_amdgcn_readfirstlane(ph[c0 + (i < w ? i : 0)]);
    // U rows: u_i -= l_ik u_k for k ascending (each entry's FMA sequence as in the unblocked
    // loop), one column per lane; L11 from the panel's compact column-major tile, where column
    // k's entries below the diagonal are contiguous (vector loads, same address on all lanes)
    const TT* Lc = panel_l11<TT>(W, c0);
    for (int t = lane; t < m; t += 64) {
        const int j = ja + g + gs * t;
        TT u[kPanel];
#pragma unroll
        for (int i = 0; i < kPanel; ++i) u[i] = A[pr[i] * lda + j];
#pragma unroll
        for (int i = 0; i < kPanel; ++i) pin(u[i]);
        u_rows_group<TT, 0>(Lc, u);
        u_rows_group<TT, 4>(Lc, u);
        u_rows_group<TT, 8>(Lc, u);
        u_rows_group<TT, 12>(Lc, u);
#pragma unroll
        for (int i = 1; i < kPanel; ++i)
            if (kFull || i < w) A[pr[i] * lda + j] = u[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // trailing rows
    TT L[NSA][kPanel];
    int prw[NSA];
    bool ok[NSA];
#pragma unroll
    for (int s = 0; s < NSA; ++s) {
        const int li = j0 + lane + 64 * s;
        ok[s] = li < n;
        prw[s] = ph[ok[s] ? li : 0];
#pragma unroll
        for (int i = 0; i < kPanel; ++i) L[s][i] = A[prw[s] * lda + c0 + (i < w ? i : 0)];
#pragma unroll
        for (int i = 0; i < kPanel; ++i) pin(L[s][i]);
#pragma unroll
        for (int i = 0; i < kPanel; ++i) L[s][i] = (ok[s] && (kFull || i < w)) ? L[s][i] : TT(0);
    }
    for (int t = 0; t < m; t += 2) {
        const int j = ja + g + gs * t;
        const bool two = t + 1 < m;
        const int jq = two ? j + gs : j;
        TT u0[kPanel], u1[kPanel];
#pragma unroll
        for (int i = 0; i < kPanel; ++i) {
            u0[i] = A[pr[i] * lda + j];
            u1[i] = A[pr[i] * lda + jq];
        }
        if (!kFull) {
#pragma unroll
            for (int i = 0; i < kPanel; ++i) {
                pin(u0[i]);
                pin(u1[i]);
            }
#pragma unroll
            for (int i = 0; i < kPanel; ++i) {
                u0[i] = i < w ? u0[i] : TT(0);
                u1[i] = i < w ? u1[i] : TT(0);
            }
        }
#pragma unroll
        for (int s = 0; s < NSA; ++s) {
            if (ok[s]) {
                TT* Ar = A + prw[s] * lda;
                TT a0 = Ar[j];
                TT a1 = Ar[jq];
#pragma unroll
                for (int i = 0; i < kPanel; ++i) {
                    a0 -= L[s][i] * u0[i];
                    a1 -= L[s][i] * u1[i];
                }
                Ar[j] = a0;
                if (two) Ar[jq] = a1;
            }
        }
    }
}

template <typename TT>
__device__ __forceinline__ void lu_apply(TT* A, int n, int lda, int c0, int w, int ja, int jb, int g, int gs,
                                         const LuScratch& W) {
    // NSA lane <-> row slots for the n - (c0 + w) <= 127 trailing rows
    if (n - (c0 + w) > 64) {
        if (w == kPanel)
            lu_apply_panel<TT, true, 2>(A, n, lda, c0, w, ja, jb, g, gs, W);
        else
            lu_apply_panel<TT, false, 2>(A, n, lda, c0, w, ja, jb, g, gs, W);
    } else if (w == kPanel) {
        lu_apply_panel<TT, true, 1>(A, n, lda, c0, w, ja, jb, g, gs, W);
    } else {
        lu_apply_panel<TT, false, 1>(A, n, lda, c0, w, ja, jb, g, gs, W);
    }
}

// Panels with look-ahead: all waves first apply panel p to their share of the next panel's
// columns; the waves other than wave 0 count themselves done on an LDS counter, and wave 0
// waits for the count (every other wave adds one per panel unconditionally, so the wait ends)
// and factors the next panel while the others apply panel p to the remaining columns.  One
// barrier per panel.  The other waves read row maps that wave 0 rewrites meanwhile only where
// they do not change (panel p's pivot rows; the set of trailing rows, taken from panel p's own
// mapping buffer).
template <typename TT>
__device__ __forceinline__ void lu_factor_augmented(TT* A, int n, int lda, int* piv, const LuScratch& W,
                                                    int nrhs = 1) {
    static_assert(kThreads >= 128, "the blocked LU needs wave 0 and at least one more wave");
    const int tid = threadIdx.x, wv = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int nc = n + nrhs;  // right-hand sides ride along as columns n .. n + nrhs - 1
    for (int r = tid; r < kLuTile; r += blockDim.x) {
        W.phys_of[r] = r;
        W.phys_of[kLuTile + r] = r;
    }
    if (tid == 0) *W.sync = 0;
    __syncthreads();
    if (wv == 0) lu_panel<TT>(A, n, lda, 0, n < kPanel ? n : kPanel, piv, W);
    __syncthreads();
    int done = 0;  // look-ahead count wave 0 waits for
    for (int c0 = 0; c0 < n; c0 += kPanel) {
        const int w = (n - c0) < kPanel ? (n - c0) : kPanel;
        const int j0 = c0 + w;
        const int wn = j0 < n ? ((n - j0) < kPanel ? (n - j0) : kPanel) : 0;  // next panel's width
        KYN_SPAN_BEGIN(t_sp);
        if (wv == 0) {
            if (wn > 0) {
                lu_apply<TT>(A, n, lda, c0, w, j0, j0 + wn, 0, nw, W);  // wave 0's share of the look-ahead
                done += nw - 1;
                // bounded wait (every other wave adds one per panel, so it ends; the bound keeps a
                // broken count from hanging the kernel: the wave then goes on and flags the error)
                for (int polls = 0;
                     __hip_atomic_load(W.sync, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < done; ++polls) {
                    if (W.wait_limit < 0 || polls >= W.wait_limit) {
                        if ((tid & 63) == 0)
                            __hip_atomic_fetch_or(W.error, kDevErrLuWait, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                KYN_SPAN_END(0, t_sp, 0);
                lu_panel<TT>(A, n, lda, j0, wn, piv, W);
            }
            KYN_SPAN_END(1, t_sp, 0);
        } else {
            if (wn > 0) {
                lu_apply<TT>(A, n, lda, c0, w, j0, j0 + wn, wv, nw, W);
                if ((tid & 63) == 0)
                    __hip_atomic_fetch_add(W.sync, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            lu_apply<TT>(A, n, lda, c0, w, j0 + wn, nc, wv - 1, nw - 1, W);
            KYN_SPAN_END(2, t_sp, 64);
        }
        __syncthreads();
        KYN_SPAN_END(3, t_sp, 0);
    }
    if (((n - 1) / kPanel) & 1) {  // final mapping into buffer 0
        for (int i = tid; i < n; i += blockDim.x) W.phys_of[i] = W.phys_of[kLuTile + i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Unpivoted blocked LU of the FP32 tangent (blocks flagged block_nopiv at setup: dynamic
// solve, every node of the block on a beam element, no translations-only nodes).  There the
// conditioned iteration matrix beta h^2 (beta' M + gamma' G + K T) of the node rows is a small
// perturbation of an SPD matrix (mass plus stiffness), which factors stably without row
// interchanges, and the constraint rows come last: their pivots are those of the Schur
// complement -B (beta h^2 K)^-1 B^T, after which nothing else is eliminated.  (The FP64 forms
// keep the pivoted LU above, whose factors are the oracle's bit for bit.)
//
// Panels of 16 columns, two barriers each:
//  P1  every wave factors the 16 x 16 diagonal tile in lanes 0..15 (row per lane, each pivot
//      row to the wave by readlane, the pivot's hardware reciprocal: no cross-wave traffic)
//      and applies it, as it goes, to up to 48 rows below in lanes 16..63 (L21 = A21 U11^-1,
//      rows split over the waves);
//  P2  in the same wave, with no barrier: U12 = L11^-1 A12 for 48 columns per wave in lanes
//      16..63 (right-hand side column n included), L11 broadcast from a column-major LDS copy
//      of the factored tile that every wave writes identically; then the barrier, after which
//      wave 0 writes the tile back;
//  P3  A22 -= L21 U12 on the matrix cores: 16 x 16 output tiles, each the sum of four
//      v_mfma_f32_16x16x4f32 products over the panel's 16 columns, tiles round-robin over the
//      waves, rows past n - 1 and columns past n masked.
// FP32 arithmetic with explicit FMAs (no reliance on contraction).  Afterwards column n holds
// y = L^-1 b and the row map is the identity.
// ---------------------------------------------------------------------------------------
template <int J>
__device__ __forceinline__ void np_tile_steps(float (&a)[kPanel], int w, int lane) {
    if constexpr (J < kPanel) {
        if (J < w) {  // uniform
            float pr[kPanel];
#pragma unroll
            for (int c = J; c < kPanel; ++c) pr[c] = rdlane(a[c], J);  // pivot row J (lane J)
            const float inv = __builtin_amdgcn_rcpf(pr[J]);  // hardware reciprocal (1 ulp): FP64 refinement follows
            if (lane > J) {  // tile rows below the pivot and every L21 row
                const float l = a[J] * inv;
                a[J] = l;
#pragma unroll
                for (int c = J + 1; c < kPanel; ++c) a[c] = __builtin_fmaf(-l, pr[c], a[c]);
            }
        }
        np_tile_steps<J + 1>(a, w, lane);
    }
}

__device__ __forceinline__ void lu_factor_nopiv_f32(float* A, int n, int lda, const LuScratch& W) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nw = blockDim.x >> 6;
    for (int r = tid; r < kLuTile; r += blockDim.x) W.phys_of[r] = r;
    for (int k0 = 0; k0 < n; k0 += kPanel) {
        KYN_SPAN_BEGIN(t_p1);
        const int w = (n - k0) < kPanel ? (n - k0) : kPanel;
        const int r0 = k0 + kPanel;  // first row / column after the panel
        // ---- P1: diagonal tile + this wave's rows below.  Every wave reads the tile, so it is
        // written back only after the barrier (by wave 0, during P3, which does not touch it);
        // the L21 rows belong to one wave each and go back at once.
        float a[kPanel];
        {
            const int row = lane < kPanel ? k0 + lane : r0 + wv * (64 - kPanel) + (lane - kPanel);
            const bool ok = lane < kPanel ? lane < w : row < n;
            const int rs = ok ? row : k0;  // clamped: loads unconditional, values selected after
            const float* ar = A + rs * lda + k0;
            if (w == kPanel) {  // uniform: every panel but a partial last one, constant offsets
#pragma unroll
                for (int c = 0; c < kPanel; ++c) a[c] = ar[c];
#pragma unroll
                for (int c = 0; c < kPanel; ++c) pin(a[c]);  // after all loads: one wait
#pragma unroll
                for (int c = 0; c < kPanel; ++c) a[c] = ok ? a[c] : 0.f;
            } else {
#pragma unroll
                for (int c = 0; c < kPanel; ++c) a[c] = ar[c < w ? c : 0];
#pragma unroll
                for (int c = 0; c < kPanel; ++c) pin(a[c]);
#pragma unroll
                for (int c = 0; c < kPanel; ++c) a[c] = (ok && c < w) ? a[c] : 0.f;
            }
            np_tile_steps<0>(a, w, lane);
            if (ok && lane >= kPanel) {
#pragma unroll
                for (int c = 0; c < kPanel; ++c) A[row * lda + k0 + c] = a[c];
            }
            // the factored tile, column-major, for P2's broadcast reads: every wave computed the
            // same tile, so every wave writes the same values here and reads its own writes
            if (lane < kPanel) {
                float* Lt = panel_l11<float>(W, k0);
#pragma unroll
                for (int c = 0; c < kPanel; ++c) Lt[c * kPanel + lane] = a[c];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        KYN_SPAN_END(7, t_p1, 0);
        KYN_SPAN_BEGIN(t_p2);
        // ---- P2, before the barrier: U12 = L11^-1 A12 for the columns after the panel (the rhs
        // column n included; a partial last panel has only the rhs after it), 48 columns per wave
        // in lanes 16..63.  L11 comes from the column-major tile written at the end of P1 (16-byte
        // loads, one address for all lanes).  These columns are disjoint from every entry P1
        // reads or writes in any wave.
        const float* Lt = panel_l11<float>(W, k0);
        for (int jb = k0 + w + wv * (64 - kPanel); jb <= n; jb += nw * (64 - kPanel)) {
            const int j = jb + lane - kPanel;
            const bool okj = lane >= kPanel && j <= n;
            const int js = okj ? j : k0;  // clamped: loads unconditional
            float u[kPanel];
#pragma unroll
            for (int i = 0; i < kPanel; ++i) u[i] = A[(k0 + (i < w ? i : 0)) * lda + js];
#pragma unroll
            for (int i = 0; i < kPanel; ++i) pin(u[i]);
#pragma unroll
            for (int i = 0; i < kPanel; ++i) u[i] = (okj && i < w) ? u[i] : 0.f;
            u_rows_group_fma<0>(Lt, u);
            u_rows_group_fma<4>(Lt, u);
            u_rows_group_fma<8>(Lt, u);
            u_rows_group_fma<12>(Lt, u);
            if (okj) {
#pragma unroll
                for (int i = 1; i < kPanel; ++i)
                    if (i < w) A[(k0 + i) * lda + j] = u[i];
            }
        }
        __syncthreads();
        KYN_SPAN_END(8, t_p2, 0);
        KYN_SPAN_BEGIN(t_p3);
        if (wv == 0 && lane < w) {  // the factored tile: U11 and the unit L11 below its diagonal
#pragma unroll
            for (int c = 0; c < kPanel; ++c)
                if (c < w) A[(k0 + lane) * lda + k0 + c] = a[c];
        }
        // ---- P3: trailing update on MFMA (full panels with rows after them only)
        const int m = n - r0;       // trailing rows
        const int mc = n + 1 - r0;  // trailing columns, rhs included
        if (w == kPanel && m > 0) {
            const int TR = (m + 15) >> 4, TC = (mc + 15) >> 4;
            const int li = lane & 15, lk = lane >> 4;  // operand row/column and k within a step
            const int fr = m >> 4, fc = mc >> 4;       // tile rows / columns with no masked entry
            for (int t = wv; t < TR * TC; t += nw) {
                const int ti = t / TC, tj = t - TC * (t / TC);
                const int ia = r0 + 16 * ti + li;  // A-operand row (L21)
                const int jb = r0 + 16 * tj + li;  // B-operand / output column
                const int t2 = t + nw, ti2 = t2 / TC, tj2 = t2 - TC * (t2 / TC);
                if (ti < fr && tj < fc && t2 < TR * TC && ti2 < fr && tj2 < fc) {
                    // uniform: this wave's next tile is full too: both at once, their loads
                    // behind one wait and their two MFMA chains interleaved
                    const int ib = r0 + 16 * ti2 + li, jb2 = r0 + 16 * tj2 + li;
                    const float* pa0 = A + ia * lda + k0 + lk;
                    const float* pb0 = A + (k0 + lk) * lda + jb;
                    float* pc0 = A + (r0 + 16 * ti + 4 * lk) * lda + jb;
                    const float* pa1 = A + ib * lda + k0 + lk;
                    const float* pb1 = A + (k0 + lk) * lda + jb2;
                    float* pc1 = A + (r0 + 16 * ti2 + 4 * lk) * lda + jb2;
                    float ao0[4], bo0[4], ao1[4], bo1[4];
                    f32x4 acc0, acc1;
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        ao0[ks] = pa0[4 * ks];
                        bo0[ks] = pb0[4 * ks * lda];
                        acc0[ks] = pc0[ks * lda];
                        ao1[ks] = pa1[4 * ks];
                        bo1[ks] = pb1[4 * ks * lda];
                        acc1[ks] = pc1[ks * lda];
                    }
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {  // all 24 loads issued before the one wait
                        pin(ao0[ks]); pin(bo0[ks]); pin(ao1[ks]); pin(bo1[ks]);
                    }
                    pin(acc0); pin(acc1);
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ao0[ks], bo0[ks], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ao1[ks], bo1[ks], acc1, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        pc0[r * lda] = acc0[r];
                        pc1[r * lda] = acc1[r];
                    }
                    t = t2;
                    continue;
                }
                if (ti < fr && tj < fc) {  // uniform: a full tile, pointer steps and no guards
                    const float* pa = A + ia * lda + k0 + lk;
                    const float* pb = A + (k0 + lk) * lda + jb;
                    float* pc = A + (r0 + 16 * ti + 4 * lk) * lda + jb;
                    float ao[4], bo[4];
                    f32x4 acc;
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        ao[ks] = pa[4 * ks];
                        bo[ks] = pb[4 * ks * lda];
                        acc[ks] = pc[ks * lda];
                    }
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        pin(ao[ks]); pin(bo[ks]);
                    }
                    pin(acc);
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-ao[ks], bo[ks], acc, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) pc[r * lda] = acc[r];
                    continue;
                }
                const bool oka = ia < n, okb = jb <= n;
                float ao[4], bo[4];
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    ao[ks] = A[(oka ? ia : r0) * lda + k0 + 4 * ks + lk];
                    bo[ks] = A[(k0 + 4 * ks + lk) * lda + (okb ? jb : r0)];
                }
                f32x4 acc;
                const int ic = r0 + 16 * ti + 4 * lk;  // first output row of this lane
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = A[((okb && ic + r < n) ? ic + r : r0) * lda + (okb ? jb : r0)];
#pragma unroll
                for (int r = 0; r < 4; ++r) pin(v[r]);
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = (okb && ic + r < n) ? v[r] : 0.f;
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const float a_ = oka ? -ao[ks] : 0.f, b_ = okb ? bo[ks] : 0.f;
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a_, b_, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (okb && ic + r < n) A[(ic + r) * lda + jb] = acc[r];
            }
            __syncthreads();
        }
        KYN_SPAN_END(9, t_p3, 0);
    }
}

// x = y / d with y_inv = 1/d correctly rounded: q0 = y y_inv, e = y - d q0, q = q0 + e y_inv
// (FMA remainder correction; correctly rounded while no intermediate leaves the normal range),
// i.e. the same value as the division y / d.
__device__ __forceinline__ double div_by(double y, double d, double d_inv) {
    const double q0 = y * d_inv;
    const double e = fma(-d, q0, y);
    return fma(e, d_inv, q0);
}
// Back substitution U x = y with y in column n of the factored augmented matrix (logical row
// i at physical row phys_of[i]), in blocks of kBackBlock rows from the bottom.  Once a block's
// x is known, wave 0 updates the rows of the block above with it and solves that block (lane <->
// row, x_k to the block's lanes by readlane), while the other waves update all rows further up
// (thread <-> row); one barrier per block.  Each row receives r_i -= u_ik x_k for k descending
// and x_k = r_k / u_kk exactly as in the column-oriented loop, so x is the same, bit for bit.
constexpr int kBackBlock = 16;
template <typename TT>
__device__ __forceinline__ double back_block_solve(const TT* A, int lda, const int* phys_of, int kb, int nb,
                                                   double r, int lane) {
    // rows kb .. kb + nb - 1 on lanes 0 .. nb - 1, r holding each row's value after the blocks below
    const bool mine = lane < nb;
    const int pr = phys_of[kb + (mine ? lane : 0)];
    const double d = mine ? (double)A[pr * lda + kb + lane] : 1.;
    const double dinv = 1. / d;
    TT u[kBackBlock];
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) u[c] = A[pr * lda + kb + (c < nb ? c : 0)];
#pragma unroll
    for (int c = kBackBlock - 1; c >= 0; --c) {
        if (c < nb) {  // uniform
            // FP64 factors: the correctly rounded quotient (the oracle's division); FP32 factors:
            // the product with the reciprocal (one rounding more, far below the factors' own error)
            const double xk = rdlane(sizeof(TT) == 8 ? div_by(r, d, dinv) : r * dinv, c);
            const double upd = fma(-(double)u[c], xk, r);
            r = lane < c ? upd : (lane == c ? xk : r);
        }
    }
    return r;
}
// row i -= its entries in block columns kb .. kb + nb - 1 times x, columns descending
template <typename TT>
__device__ __forceinline__ double back_row_update(const TT* A, int lda, const int* phys_of, int i, int kb, int nb,
                                                  const double* xs, double r) {
    const TT* Ar = A + phys_of[i] * lda + kb;
    TT u[kBackBlock];
    double xk[kBackBlock];
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) {  // loads first (clamped to the block), one wait
        u[c] = Ar[c < nb ? c : 0];
        xk[c] = xs[kb + (c < nb ? c : 0)];
    }
#pragma unroll
    for (int c = kBackBlock - 1; c >= 0; --c) r = c < nb ? fma(-(double)u[c], xk[c], r) : r;
    return r;
}
template <typename TT>
__device__ __forceinline__ void back_substitute(const TT* A, int n, int lda, const int* phys_of, double* xs,
                                                bool y_in_matrix = true) {
    const int tid = threadIdx.x, lane = tid & 63;
    if (y_in_matrix) {  // else the caller has written y (FP64, logical order) to xs and synchronized
        for (int i = tid; i < n; i += blockDim.x) xs[i] = (double)A[phys_of[i] * lda + n];
        __syncthreads();
    }
    int kb = ((n - 1) / kBackBlock) * kBackBlock;
    int nb = n - kb;  // the bottom block, possibly partial
    if (tid < 64) {
        const double r = back_block_solve<TT>(A, lda, phys_of, kb, nb, lane < nb ? xs[kb + lane] : 0., lane);
        if (lane < nb) xs[kb + lane] = r;
    }
    __syncthreads();
    for (; kb > 0; kb -= kBackBlock) {
        const int kp = kb - kBackBlock;  // the block above (full)
        if (tid < 64) {
            double r = 0.;
            if (lane < kBackBlock) r = back_row_update<TT>(A, lda, phys_of, kp + lane, kb, nb, xs, xs[kp + lane]);
            r = back_block_solve<TT>(A, lda, phys_of, kp, kBackBlock, r, lane);
            if (lane < kBackBlock) xs[kp + lane] = r;
        } else {
            for (int i = tid - 64; i < kp; i += blockDim.x - 64)
                xs[i] = back_row_update<TT>(A, lda, phys_of, i, kb, nb, xs, xs[i]);
        }
        __syncthreads();
        nb = kBackBlock;
    }
}

// Back substitution U X = Y for the right-hand sides in columns n + c, c in [0, ncols), of the
// factored augmented matrix, all columns together in blocks of 16 rows from the bottom: per
// block, each wave solves the diagonal block for four columns (16-lane group <-> column, lane
// <-> row, x_j to its group by a lane shuffle), then every thread updates one (row, column) of
// the rows above with the block's x.  Each entry receives r_i -= u_ij x_j for j descending and
// x_j = r_j / u_jj (the IEEE quotient) as in the column-oriented loop.  X: LDS, [ncols][n]
// doubles, the solution on return; also written to out[c * ldo + i] (HBM).
template <typename TT>
__device__ __forceinline__ void back_substitute_cols_blocked(const TT* A, int n, int lda, const int* phys_of,
                                                             int ncols, double* X, double* out, int ldo) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    for (int q = tid; q < ncols * n; q += blockDim.x) {
        const int c = q / n, i = q - n * (q / n);
        X[q] = (double)A[phys_of[i] * lda + n + c];
    }
    __syncthreads();
    const int g = lane >> 4, rl = lane & 15;
    for (int kb = ((n - 1) / 16) * 16; kb >= 0; kb -= 16) {
        const int nb = (n - kb) < 16 ? (n - kb) : 16;
        for (int cb = 4 * wv; cb < ncols; cb += 4 * nw) {  // the same for every lane of a wave
            const int c = cb + g;
            const bool okc = c < ncols, okr = rl < nb;
            const int i = kb + (okr ? rl : 0);
            const int pr = phys_of[i];
            double r = okc && okr ? X[c * n + i] : 0.;
            const double d = okr ? (double)A[pr * lda + i] : 1.;
            const double dinv = 1. / d;
            TT u[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) u[j] = A[pr * lda + kb + (j < nb ? j : 0)];
#pragma unroll
            for (int j = 15; j >= 0; --j) {
                if (j < nb) {  // uniform
                    const double xj = __shfl(div_by(r, d, dinv), (lane & ~15) | j, 64);
                    const double upd = fma(-(double)u[j], xj, r);
                    r = rl < j ? upd : (rl == j ? xj : r);
                }
            }
            if (okc && okr) X[c * n + i] = r;
        }
        __syncthreads();
        for (int q = tid; q < kb * ncols; q += blockDim.x) {
            const int c = q / kb, i = q - kb * (q / kb);
            const TT* Ar = A + phys_of[i] * lda + kb;
            const double* xb = X + c * n + kb;
            double r = X[c * n + i];
            for (int j = nb - 1; j >= 0; --j) r = fma(-(double)Ar[j], xb[j], r);
            X[c * n + i] = r;
        }
        __syncthreads();
    }
    for (int q = tid; q < ncols * n; q += blockDim.x) {
        const int c = q / n, i = q - n * (q / n);
        out[(size_t)c * ldo + i] = X[q];
    }
}

// Forward substitution L y = r (unit L; logical row i at physical row phys_of[i]) on xs (FP64,
// logical order, in place), in blocks of 16 rows from the top, mirroring back_substitute: once a
// block's y is known, wave 0 updates the rows of the next block with it and solves that block
// (lane <-> row, y_k to the block's lanes by readlane), while the other waves update all rows
// further down (thread <-> row); one barrier per block.  Each row receives r_i -= l_ik y_k for
// k ascending, as in the column-oriented loop.
template <typename TT>
__device__ __forceinline__ double fwd_block_solve(const TT* A, int lda, const int* phys_of, int kb, int nb, double r,
                                                  int lane) {
    const bool mine = lane < nb;
    const int pr = phys_of[kb + (mine ? lane : 0)];
    TT l[kBackBlock];
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) l[c] = A[pr * lda + kb + (c < nb ? c : 0)];
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) {
        if (c < nb) {  // uniform
            const double yk = rdlane(r, c);
            const double upd = fma(-(double)l[c], yk, r);
            r = lane > c ? upd : r;
        }
    }
    return r;
}
// row i -= its entries in block columns kb .. kb + 15 times y, columns ascending
template <typename TT>
__device__ __forceinline__ double fwd_row_update(const TT* A, int lda, const int* phys_of, int i, int kb,
                                                 const double* xs, double r) {
    const TT* Ar = A + phys_of[i] * lda + kb;
    TT l[kBackBlock];
    double yk[kBackBlock];
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) {  // loads first, one wait
        l[c] = Ar[c];
        yk[c] = xs[kb + c];
    }
#pragma unroll
    for (int c = 0; c < kBackBlock; ++c) r = fma(-(double)l[c], yk[c], r);
    return r;
}
template <typename TT>
__device__ __forceinline__ void forward_substitute(const TT* A, int n, int lda, const int* phys_of, double* xs) {
    const int tid = threadIdx.x, lane = tid & 63;
    int nb = n < kBackBlock ? n : kBackBlock;
    if (tid < 64) {
        const double r = fwd_block_solve<TT>(A, lda, phys_of, 0, nb, lane < nb ? xs[lane] : 0., lane);
        if (lane < nb) xs[lane] = r;
    }
    __syncthreads();
    for (int kb = 0; kb + kBackBlock < n; kb += kBackBlock) {  // block kb is solved: apply it below
        const int kn = kb + kBackBlock;
        const int nbn = (n - kn) < kBackBlock ? (n - kn) : kBackBlock;
        if (tid < 64) {
            double r = 0.;
            if (lane < nbn) r = fwd_row_update<TT>(A, lda, phys_of, kn + lane, kb, xs, xs[kn + lane]);
            r = fwd_block_solve<TT>(A, lda, phys_of, kn, nbn, r, lane);
            if (lane < nbn) xs[kn + lane] = r;
        } else {
            for (int i = kn + kBackBlock + tid - 64; i < n; i += blockDim.x - 64)
                xs[i] = fwd_row_update<TT>(A, lda, phys_of, i, kb, xs, xs[i]);
        }
        __syncthreads();
    }
}

// Triangular solves with the LDS factors (logical row i at physical row phys_of[i]), FP64
// accumulation, both blocked over all waves.  b: LDS, original row order, read through phys_of
// (not modified); x -> xs.
template <typename TT>
__device__ __forceinline__ void lu_solve(const TT* A, int n, int lda, const int* phys_of, const double* b, double* xs) {
    KYN_SPAN_BEGIN(t_fw);
    // P b: the factorization's row interchanges (k <-> piv[k], k ascending) composed on the
    // identity are phys_of, so entry i of the interchanged b is b[phys_of[i]]: one gather
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = b[phys_of[i]];
    __syncthreads();
    forward_substitute<TT>(A, n, lda, phys_of, xs);
    KYN_SPAN_END(6, t_fw, 0);
    back_substitute<TT>(A, n, lda, phys_of, xs, false);
}

// ---------------------------------------------------------------------------------------
// Node-level kernels shared by both forms
// ---------------------------------------------------------------------------------------
// PredictNextState (src/step/predict_next_state.hpp:14-47, calculate_next_state.hpp:26-37)
__device__ __forceinline__ void predict_node(const DevSys& S, int node) {
    const StepConsts& p = S.p;
    // every load first (the state arrays may alias as far as the compiler knows, so loads
    // interleaved with stores would each wait for HBM), then the arithmetic, then the stores
    double qp[7], v_p[6], vd_p[6], a_p[6];
#pragma unroll
    for (int c = 0; c < 7; ++c) qp[c] = S.q[node * 7 + c];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        v_p[c] = S.v[node * 6 + c];
        vd_p[c] = S.vd[node * 6 + c];
        a_p[c] = S.a[node * 6 + c];
    }
    double an[6], vn[6], qd[6], qn[7];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        an[c] = (p.alpha_f * vd_p[c] - p.alpha_m * a_p[c]) / (1. - p.alpha_m);
        vn[c] = v_p[c] + p.h * (1. - p.gamma) * a_p[c] + p.gamma * p.h * an[c];
        qd[c] = v_p[c] + (0.5 - p.beta) * p.h * a_p[c] + p.beta * p.h * an[c];
    }
    calc_displacement(p.h, qd, qp, qn);
#pragma unroll
    for (int c = 0; c < 7; ++c) S.q_prev[node * 7 + c] = qp[c];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const int k = node * 6 + c;
        S.vd[k] = 0.;
        S.a[k] = an[c];
        S.v[k] = vn[c];
        S.q_delta[k] = qd[c];
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) S.q[node * 7 + c] = qn[c];
}

__device__ __forceinline__ void displacement_node(const DevSys& S, int node) {
    double qd[6], qp[7], qn[7];
#pragma unroll
    for (int c = 0; c < 6; ++c) qd[c] = S.q_delta[node * 6 + c];
#pragma unroll
    for (int c = 0; c < 7; ++c) qp[c] = S.q_prev[node * 7 + c];
    calc_displacement(S.p.h, qd, qp, qn);
#pragma unroll
    for (int c = 0; c < 7; ++c) S.q[node * 7 + c] = qn[c];
}

// Step epilogue (src/step/step.hpp:85-121)
__device__ __forceinline__ void epilogue_node(const DevSys& S, int node) {
    const double f = (1. - S.p.alpha_f) / (1. - S.p.alpha_m);
#pragma unroll
    for (int c = 0; c < 6; ++c) S.a[node * 6 + c] += f * S.vd[node * 6 + c];
    const double* q = S.q + node * 7;
    const double* x0 = S.x0 + node * 7;
    double* x = S.x + node * 7;
    x[0] = x0[0] + q[0];
    x[1] = x0[1] + q[1];
    x[2] = x0[2] + q[2];
    const double a[4] = {q[3], q[4], q[5], q[6]}, b[4] = {x0[3], x0[4], x0[5], x0[6]};
    double r[4];
    quat_compose(a, b, r);
    x[3] = r[0];
    x[4] = r[1];
    x[5] = r[2];
    x[6] = r[3];
}

// calculate_revolute_joint_output.hpp:30-86
__device__ __forceinline__ void constraint_output(const DevSys& S, int c) {
    if (S.con_type[c] != 4) return;
    const int node = S.con_target[c];
    const double* q = S.q + node * 7;
    const double R[4] = {q[3], q[4], q[5], q[6]};
    const double theta = 2. * acos(R[0]);
    const double shh = sqrt(1. - R[0] * R[0]);
    double rv[3] = {0., 0., 0.};
    if (shh > 1e-12) {
        rv[0] = theta * R[1] / shh;
        rv[1] = theta * R[2] / shh;
        rv[2] = theta * R[3] / shh;
    }
    double ax[3];
    quat_rotate(R, S.con_axes + c * 9, ax);
    const double* v = S.v + node * 6;
    const double* vd = S.vd + node * 6;
    S.con_output[c * 3 + 0] = ax[0] * rv[0] + ax[1] * rv[1] + ax[2] * rv[2];
    S.con_output[c * 3 + 1] = ax[0] * v[3] + ax[1] * v[4] + ax[2] * v[5];
    S.con_output[c * 3 + 2] = ax[0] * vd[3] + ax[1] * vd[4] + ax[2] * vd[5];
}

// Constraint gradient times stored tangent: G T with T = [I 0; 0 Trr]
__device__ __forceinline__ void times_tangent(const double* g, const double* trr, double* out) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 6 + j] = g[i * 6 + j];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            out[i * 6 + 3 + j] = g[i * 6 + 3] * trr[j] + g[i * 6 + 4] * trr[3 + j] + g[i * 6 + 5] * trr[6 + j];
    }
}

// ---------------------------------------------------------------------------------------
// Schedule application (device-resident per-step inputs)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void apply_schedule_model(const DevSys& S, int m) {
    if (S.sched_steps <= 0) return;
    const int slot = *S.sched_slot;
    if (slot >= S.sched_steps) return;
    if (S.sched_cdisp)
        for (int c = S.model_con_begin[m] + (int)threadIdx.x; c < S.model_con_end[m]; c += blockDim.x)
#pragma unroll
            for (int k = 0; k < 7; ++k)
                S.con_input[c * 7 + k] = S.sched_cdisp[((size_t)slot * S.num_cons + c) * 7 + k];
    if (S.sched_loads)
        for (int k = S.model_load_begin[m] + (int)threadIdx.x; k < S.model_load_begin[m + 1]; k += blockDim.x) {
            const int node = S.sched_load_nodes[k];
#pragma unroll
            for (int c = 0; c < 6; ++c)
                S.f[node * 6 + c] = S.sched_loads[((size_t)slot * S.sched_num_loaded + k) * 6 + c];
        }
}

__global__ void k_advance_slot(DevSys S) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && S.sched_steps > 0) *S.sched_slot += 1;
}

// ---------------------------------------------------------------------------------------
// Block phase helpers (fused form): constraints into the LDS KKT block
// ---------------------------------------------------------------------------------------
// Translations-only nodes (FreedomSignature::JustPosition: springs or three-dof constraints,
// three dofs in the reference): the block keeps six rows per node, and a pinned node's three
// rotational rows and columns hold nothing but a unit diagonal and a zero rhs, so their
// increments are exactly zero and the elimination of every other row is unchanged.
template <typename TT>
__device__ __forceinline__ void pin_rotations(const DevSys& S, int b, TT* A, int lda, double* rhs, int npin = -1) {
    if ((npin < 0 ? S.block_npin[b] : npin) == 0) return;  // uniform over the workgroup
    for (int k = S.block_node_begin[b] + threadIdx.x; k < S.block_node_end[b]; k += blockDim.x) {
        const int node = S.block_nodes[k];
        if (!S.node_pin[node]) continue;
        const int l = S.node_local[node];
        for (int c = 3; c < 6; ++c) {
            A[(size_t)(l + c) * lda + l + c] = TT(1);
            rhs[l + c] = 0.;
        }
    }
    __syncthreads();
}

template <typename TT>
__device__ __forceinline__ void block_constraints(const DevSys& S, int b, char* sm, const Lay& L) {
    TT* A = reinterpret_cast<TT*>(sm + L.A);
    double* cres = reinterpret_cast<double*>(sm + L.cres);
    const int c0 = S.block_con_begin[b], c1 = S.block_con_end[b];
    const int nc = c1 - c0;
    const int t = threadIdx.x;
    // One constraint (every fused blade block): lane 0 of wave 1 forms its target's rotational
    // tangent block and lane 0 of wave 2 its base's, into the reduction scratch (free in this
    // phase), while thread 0 evaluates the constraint; one barrier joins them.  Otherwise thread
    // c does all of constraint c.
    const bool split = nc == 1;  // uniform
    const int npin = S.block_npin[b];  // loaded now, used after the last barrier
    double* trr_sh = reinterpret_cast<double*>(sm + L.red);  // [2][9]: target, base
    if (split && (t == 64 || t == 128)) {
        const int c = S.block_cons[c0];
        const bool base = t == 128;
        if (!base || con_two_node(S.con_type[c])) {
            const int nd = base ? S.con_base[c] : S.con_target[c];
            double qd[6], trr[9];
#pragma unroll
            for (int k = 0; k < 6; ++k) qd[k] = S.q_delta[nd * 6 + k];
            tangent_rot_block(S.p.h, qd, trr);
#pragma unroll
            for (int k = 0; k < 9; ++k) trr_sh[(base ? 9 : 0) + k] = trr[k];
        }
    }
    int c = 0, type = 0, tn = 0, bn = 0, rows = 0, lr = 0, lt = 0, lb = 0;
    bool two = false;
    double res[6], bg[36], tg[36], sres[6] = {0., 0., 0., 0., 0., 0.}, lam[6];
    if (t < nc) {
        c = S.block_cons[c0 + t];
        type = S.con_type[c];
        two = con_two_node(type);
        tn = S.con_target[c];
        bn = S.con_base[c];
        rows = S.con_rows[c];
        lr = S.con_local[c];
        lt = S.node_local[tn];
        lb = two ? S.node_local[bn] : 0;
        double input[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) input[k] = S.con_input[c * 7 + k];
        if (con_controlled(type)) input[0] = S.con_control[c];  // Constraints::UpdateViews
#pragma unroll
        for (int k = 0; k < 6; ++k) lam[k] = S.con_lambda[c * 6 + k];
        double tq[7], bq[7] = {0., 0., 0., 1., 0., 0., 0.};
#pragma unroll
        for (int k = 0; k < 7; ++k) tq[k] = S.q[tn * 7 + k];
        if (two)
#pragma unroll
            for (int k = 0; k < 7; ++k) bq[k] = S.q[bn * 7 + k];
        constraint_eval(type, S.con_X0 + c * 3, S.con_axes + c * 9, input, bq, tq, res, bg, tg, sres);
    }
    if (split) __syncthreads();
    if (t < nc) {
        double trr[9], qd[6], gt[36];
        if (split) {
#pragma unroll
            for (int k = 0; k < 9; ++k) trr[k] = trr_sh[k];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) qd[k] = S.q_delta[tn * 6 + k];
            tangent_rot_block(S.p.h, qd, trr);
        }
        times_tangent(tg, trr, gt);
        double* cr = cres + t * 24;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            if (i < rows)
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    A[(lr + i) * L.lda + lt + j] = (TT)gt[i * 6 + j];
                    A[(lt + j) * L.lda + lr + i] = (TT)tg[i * 6 + j];
                }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double acc = 0.;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += tg[k * 6 + i] * lam[k];
            cr[12 + i] = acc;  // target lambda residual
        }
        if (two) {
            if (split) {
#pragma unroll
                for (int k = 0; k < 9; ++k) trr[k] = trr_sh[9 + k];
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) qd[k] = S.q_delta[bn * 6 + k];
                tangent_rot_block(S.p.h, qd, trr);
            }
            times_tangent(bg, trr, gt);
#pragma unroll
            for (int i = 0; i < 6; ++i)
                if (i < rows)
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        A[(lr + i) * L.lda + lb + j] = (TT)gt[i * 6 + j];
                        A[(lb + j) * L.lda + lr + i] = (TT)bg[i * 6 + j];
                    }
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double acc = 0.;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc += bg[k * 6 + i] * lam[k];
                cr[6 + i] = acc;
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            cr[k] = res[k];
            cr[18 + k] = sres[k];
            S.con_res[c * 6 + k] = res[k];
            S.con_sys_res[c * 6 + k] = sres[k];
        }
        if (split) {  // the one constraint's rhs terms, in the order of the loop below
            double* rhs = reinterpret_cast<double*>(sm + L.rhs);
            for (int j = 0; j < 6; ++j) rhs[lt + j] += cr[18 + j];
            if (two)
                for (int j = 0; j < 6; ++j) rhs[lb + j] += cr[6 + j];
            for (int j = 0; j < 6; ++j) rhs[lt + j] += cr[12 + j];
            for (int i = 0; i < rows; ++i) rhs[lr + i] = 0. + cr[i];
        }
    }
    __syncthreads();
    // rhs in the reference's order (assemble_constraints_residual.hpp:27-50)
    if (t == 0 && nc > 0 && !split) {
        double* rhs = reinterpret_cast<double*>(sm + L.rhs);
        for (int k = 0; k < nc; ++k) {
            const int c = S.block_cons[c0 + k];
            const int lt = S.node_local[S.con_target[c]];
            for (int j = 0; j < 6; ++j) rhs[lt + j] += cres[k * 24 + 18 + j];
        }
        for (int k = 0; k < nc; ++k) {
            const int c = S.block_cons[c0 + k];
            const int type = S.con_type[c];
            if (con_two_node(type)) {
                const int lb = S.node_local[S.con_base[c]];
                for (int j = 0; j < 6; ++j) rhs[lb + j] += cres[k * 24 + 6 + j];
            }
            const int lt = S.node_local[S.con_target[c]];
            for (int j = 0; j < 6; ++j) rhs[lt + j] += cres[k * 24 + 12 + j];
        }
        for (int k = 0; k < nc; ++k) {
            const int c = S.block_cons[c0 + k];
            const int lr = S.con_local[c];
            for (int i = 0; i < S.con_rows[c]; ++i) rhs[lr + i] = 0. + cres[k * 24 + i];
        }
    }
    __syncthreads();
    pin_rotations<TT>(S, b, A, L.lda, reinterpret_cast<double*>(sm + L.rhs), npin);
}

// Convergence-error terms and state update of one block after its solve
// (calculate_convergence_error.hpp, update_state_prediction.hpp, update_lambda_prediction.hpp)
__device__ __forceinline__ double block_error_and_update(const DevSys& S, int b, const double* xs, double* red) {
    const StepConsts& p = S.p;
    const int n = S.block_n[b], nsys = S.block_nsys[b];
    const int nb0 = S.block_node_begin[b];
    const int t = threadIdx.x;
    double term = 0.;
    if (n <= (int)blockDim.x - 64) {
        // uniform: one row per thread on the first waves, the update's loads issued with the
        // error's; the last wave forms the nodes' new displacements from the old q_delta and x
        // itself (the same q_delta + x / h the row threads store), so no barrier separates the
        // two.  Every old value is read before the reduction's barriers, every store follows them.
        const int i = t;
        double qd = 0., v = 0., vd = 0.;
        int k = 0;
        if (i < nsys) {
            k = S.block_nodes[nb0 + i / 6] * 6 + i % 6;
            qd = S.q_delta[k];
            if (p.is_dynamic) {
                v = S.v[k];
                vd = S.vd[k];
            }
        }
        // constraint multipliers: operands loaded now, sums written after the reduction
        const int qc = S.block_con_begin[b] + t;
        const bool has_con = qc < S.block_con_end[b];
        int cidx = 0, crows = 0, cloc = 0;
        double lam[6];
        if (has_con) {
            cidx = S.block_cons[qc];
            crows = S.con_rows[cidx];
            cloc = S.con_local[cidx];
#pragma unroll
            for (int r = 0; r < 6; ++r) lam[r] = r < crows ? S.con_lambda[cidx * 6 + r] : 0.;
        }
        const int nodeT0 = blockDim.x - 64;
        if (t >= nodeT0) {
            for (int kn = nb0 + t - nodeT0; kn < S.block_node_end[b]; kn += 64) {
                const int node = S.block_nodes[kn], r0 = 6 * (kn - nb0);
                double qn_d[6], qp[7], qn[7];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const double old = S.q_delta[node * 6 + c];
                    qn_d[c] = r0 + c < nsys ? old + xs[r0 + c] / p.h : old;
                }
#pragma unroll
                for (int c = 0; c < 7; ++c) qp[c] = S.q_prev[node * 7 + c];
                calc_displacement(p.h, qn_d, qp, qn);
#pragma unroll
                for (int c = 0; c < 7; ++c) S.q[node * 7 + c] = qn[c];
            }
        }
        if (i < nsys) {
            const double e = xs[i] / (p.atol + p.rtol * fabs(qd * p.h));
            term = e * e;
        } else if (i < n) {
            double e = 0.;
            for (int q = S.block_con_begin[b]; q < S.block_con_end[b]; ++q) {
                const int c = S.block_cons[q];
                const int r = i - S.con_local[c];
                if (r >= 0 && r < S.con_rows[c]) e = xs[i] / (p.atol + p.rtol * fabs(S.con_lambda[c * 6 + r]));
            }
            term = e * e;
        }
        const double sum = block_sum(term, red);
        if (i < nsys) {  // the same update as below, from the values already in registers
            const double dx = xs[i];
            S.q_delta[k] = qd + dx / p.h;
            if (p.is_dynamic) {
                S.v[k] = v + p.gamma_prime * dx;
                S.vd[k] = vd + p.beta_prime * dx;
            }
        }
        if (has_con) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r < crows) S.con_lambda[cidx * 6 + r] = lam[r] + xs[cloc + r];
        }
        for (int q = qc + blockDim.x; q < S.block_con_end[b]; q += blockDim.x) {
            const int c = S.block_cons[q];
            for (int r = 0; r < S.con_rows[c]; ++r) S.con_lambda[c * 6 + r] += xs[S.con_local[c] + r];
        }
        __syncthreads();
        return sum;
    }
    for (int i = t; i < n; i += blockDim.x) {
        double e;
        if (i < nsys) {
            const int node = S.block_nodes[nb0 + i / 6];
            const double xi = S.q_delta[node * 6 + i % 6] * p.h;
            e = xs[i] / (p.atol + p.rtol * fabs(xi));
        } else {
            // constraint row: find constraint by local row
            e = 0.;
            for (int k = S.block_con_begin[b]; k < S.block_con_end[b]; ++k) {
                const int c = S.block_cons[k];
                const int r = i - S.con_local[c];
                if (r >= 0 && r < S.con_rows[c]) {
                    e = xs[i] / (p.atol + p.rtol * fabs(S.con_lambda[c * 6 + r]));
                }
            }
        }
        term += e * e;
    }
    const double sum = block_sum(term, red);
    // update (after every thread has read q_delta / lambda for the error)
    for (int i = t; i < nsys; i += blockDim.x) {
        const int node = S.block_nodes[nb0 + i / 6];
        const int k = node * 6 + i % 6;
        const double dx = xs[i];
        S.q_delta[k] += dx / p.h;
        if (p.is_dynamic) {
            S.v[k] += p.gamma_prime * dx;
            S.vd[k] += p.beta_prime * dx;
        }
    }
    for (int k = S.block_con_begin[b] + t; k < S.block_con_end[b]; k += blockDim.x) {
        const int c = S.block_cons[k];
        for (int r = 0; r < S.con_rows[c]; ++r) S.con_lambda[c * 6 + r] += xs[S.con_local[c] + r];
    }
    __syncthreads();
    for (int k = nb0 + t; k < S.block_node_end[b]; k += blockDim.x) displacement_node(S, S.block_nodes[k]);
    __syncthreads();
    return sum;
}


// ---------------------------------------------------------------------------------------
// Conditioning (src/solver/condition_system.hpp:10-31, solve_system.hpp:20-25): keep the
// assembled b for kyn_get_rhs, scale system rows by beta*h^2, negate.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void condition_rhs(const DevSys& S, int b, double* rhs) {
    const int n = S.block_n[b], nsys = S.block_nsys[b], off = S.block_vec_off[b];
    const double cond = S.p.conditioner;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double raw = rhs[i];
        S.rhs_raw[off + i] = raw;
        const double c = i < nsys ? raw * cond : raw;
        rhs[i] = -1. * c;
        S.rhs[off + i] = rhs[i];
    }
    __syncthreads();
}

// Factor + solve of one block (replaces dss::numeric_factorization / dss::solve,
// src/step/solve_system.hpp:27-31): the conditioned rhs rides along as column n of the
// blocked LU, then back substitution; optional FP64 iterative refinement against the
// unfactored matrix (FP32 tangent; column-major copy in HBM); then UnconditionSolution
// (condition_system.hpp:24-31).  A: LDS matrix with lda >= n + 1.
template <typename TT>
__device__ __forceinline__ void factor_solve_uncondition(const DevSys& S, int b, TT* A, TT* Ac, int lda, int* piv,
                                         double* rhs, double* xs, double* xs2, char* sm, const Lay& L) {
    const int n = S.block_n[b], nsys = S.block_nsys[b], off = S.block_vec_off[b];
    const int t = threadIdx.x;
    const int ldc = (n + 3) & ~3;  // row stride of the copy: whole 16-byte rows
    if (Ac != nullptr) {  // row-major copy, zero-padded rows: the residual reads 16-byte vectors
        for (int k = t; k < n * ldc; k += blockDim.x) {
            const int i = k / ldc, j = k - ldc * (k / ldc);
            Ac[k] = j < n ? A[i * lda + j] : TT(0);
        }
    }
    for (int i = t; i < n; i += blockDim.x) A[i * lda + n] = (TT)rhs[i];
    __syncthreads();
    const LuScratch W{reinterpret_cast<int*>(sm + L.phys_of), reinterpret_cast<void*>(sm + L.lcol),
                      reinterpret_cast<int*>(sm + L.lu_sync), S.lu_wait_limit, S.dev_error};
    KYN_PHASE_MARK(7);
    if constexpr (sizeof(TT) == 4) {
        if (S.block_nopiv[b])
            lu_factor_nopiv_f32(A, n, lda, W);
        else
            lu_factor_augmented<TT>(A, n, lda, piv, W);
    } else {
        lu_factor_augmented<TT>(A, n, lda, piv, W);
    }
    KYN_PHASE_MARK(8);
    back_substitute<TT>(A, n, lda, W.phys_of, xs);
    KYN_PHASE_MARK(9);
    for (int sweep = 0; Ac != nullptr && sweep < S.refine_sweeps; ++sweep) {
        KYN_SPAN_BEGIN(t_rf);
        typedef TT vec4_t __attribute__((ext_vector_type(4)));
        for (int i = t; i < n; i += blockDim.x) {
            double r = S.rhs[off + i];
            const vec4_t* row = reinterpret_cast<const vec4_t*>(Ac + (size_t)i * ldc);
            // thread <-> row, four entries per load, j ascending
#pragma unroll 8
            for (int j4 = 0; j4 < ldc / 4; ++j4) {
                const vec4_t v = row[j4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (4 * j4 + e < n) r -= (double)v[e] * xs[4 * j4 + e];
            }
            rhs[i] = r;
        }
        __syncthreads();
        KYN_SPAN_END(4, t_rf, 0);
        KYN_SPAN_BEGIN(t_rs);
        lu_solve<TT>(A, n, lda, W.phys_of, rhs, xs2);
        KYN_SPAN_END(5, t_rs, 0);
        for (int i = t; i < n; i += blockDim.x) xs[i] += xs2[i];
        __syncthreads();
    }
    const double cond = S.p.conditioner;
    for (int i = t; i < n; i += blockDim.x) {
        if (i >= nsys) xs[i] /= cond;
        S.xsol[off + i] = xs[i];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Fused form: one workgroup per model, one launch per time step
// ---------------------------------------------------------------------------------------
template <typename TT>
__device__ __forceinline__ void fused_step_body(const DevSys& S, int Nm, int Qm, int nm) {
    extern __shared__ __align__(16) char sm[];
    const Lay L = make_layout(Nm, Qm, nm, S.max_block_cons, S.refine_sweeps > 0, (int)sizeof(TT), true);
    const int m = S.model_order[blockIdx.x];
    const int t = threadIdx.x;
    double* red = reinterpret_cast<double*>(sm + L.red);
    int* misc = reinterpret_cast<int*>(sm + L.misc);
    double* rhs = reinterpret_cast<double*>(sm + L.rhs);
    double* xs = reinterpret_cast<double*>(sm + L.xs);
    double* xs2 = reinterpret_cast<double*>(sm + L.xs2);
    int* piv = reinterpret_cast<int*>(sm + L.piv);
    TT* A = reinterpret_cast<TT*>(sm + L.A);
    const StepConsts& p = S.p;

    // inputs, PredictNextState, ResetConstraints (step.hpp:49-50)
    apply_schedule_model(S, m);
    __syncthreads();
    for (int node = S.model_node_begin[m] + t; node < S.model_node_end[m]; node += blockDim.x)
        predict_node(S, node);
    for (int c = S.model_con_begin[m] + t; c < S.model_con_end[m]; c += blockDim.x)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            S.con_lambda[c * 6 + k] = 0.;
            S.con_sys_res[c * 6 + k] = 0.;
        }
    __syncthreads();

    const double ndofs = (double)S.model_ndofs[m];
    int converged = 0;
    int iter = 0;
    double err_prev = __builtin_inf();  // convergence error of the previous Newton iteration
    for (; iter < p.max_iter; ++iter) {
        // FP64 refinement of the FP32 solve where it decides the final state: the Newton
        // increments contract by ~20x per iteration, so a solve error matters only in the last
        // iterations, whose predecessors' errors are small (an inexact-Newton forcing rule)
        const bool refine_now = S.refine_sweeps > 0 && (S.refine_err_bound <= 0. || err_prev < S.refine_err_bound);
        double err_sum = 0.;
        for (int b = S.model_block_begin[m]; b < S.model_block_end[m]; ++b) {
            const int n = S.block_n[b];
            const int e = S.block_elems[S.block_elem_begin[b]];
            KYN_PHASE_MARK(0);
            // every system row starts at -f (a node outside the element, e.g. a massless hub,
            // takes only its load and constraint terms); the element rows are then written as
            // -f + R after the element phase's first barrier
            for (int i = t; i < S.block_nsys[b]; i += blockDim.x)
                rhs[i] = -S.f[S.block_nodes[S.block_node_begin[b] + i / 6] * 6 + i % 6];
            element_phase<TT, true>(S, e, sm, L, n, nullptr, nullptr, m);
            KYN_PHASE_MARK(5);
            block_constraints<TT>(S, b, sm, L);
            condition_rhs(S, b, rhs);
            KYN_PHASE_MARK(6);
            // unfactored copy for refinement: row-major, row stride (n + 3) & ~3, at the block's own offset
            TT* Ac = (S.Acopy != nullptr && refine_now) ? reinterpret_cast<TT*>(S.Acopy) + S.block_C_off[b] : nullptr;
            factor_solve_uncondition<TT>(S, b, A, Ac, L.lda, piv, rhs, xs, xs2, sm, L);
            KYN_PHASE_MARK(10);
            err_sum += block_error_and_update(S, b, xs, red);
            KYN_PHASE_MARK(11);
        }
        const double err = sqrt(err_sum / ndofs);
        err_prev = err;
        if (t == 0) S.err_hist[(size_t)m * p.max_iter + iter] = err;
        // the reference's loop test `err > 1.` (step.hpp:54): a NaN error also ends the loop,
        // and the step then counts as converged
        if (!(err > 1.)) {
            converged = 1;
            ++iter;
            break;
        }
    }
    if (converged) {
        for (int node = S.model_node_begin[m] + t; node < S.model_node_end[m]; node += blockDim.x)
            epilogue_node(S, node);
        for (int c = S.model_con_begin[m] + t; c < S.model_con_end[m]; c += blockDim.x) constraint_output(S, c);
    }
    if (t == 0) {
        S.model_status[m] = converged ? 0 : 1;
        S.model_iters[m] = iter;
        S.model_active[m] = 0;
        S.model_iter_total[m] += iter;
        S.model_fail_total[m] += converged ? 0 : 1;
    }
}

// FP64 tangent: the KKT block alone is up to 141 KB of LDS, one workgroup per CU.
__global__ void __launch_bounds__(kThreads) k_fused_step_f64(DevSys S, int Nm, int Qm, int nm) {
    fused_step_body<double>(S, Nm, Qm, nm);
}
// FP32 tangent: blocks up to 143 dofs fit in 80 KB, so two workgroups share a CU (registers
// capped at 256 per lane for that).
__global__ void __launch_bounds__(kThreads, 2) k_fused_step_f32(DevSys S, int Nm, int Qm, int nm) {
    fused_step_body<float>(S, Nm, Qm, nm);
}

// ---------------------------------------------------------------------------------------
// Plain form: the reference's kernel sequence, one launch per stage
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_plain_predict(DevSys S) {
    const int m = blockIdx.x, t = threadIdx.x;
    apply_schedule_model(S, m);
    __syncthreads();
    for (int node = S.model_node_begin[m] + t; node < S.model_node_end[m]; node += blockDim.x)
        predict_node(S, node);
    for (int c = S.model_con_begin[m] + t; c < S.model_con_end[m]; c += blockDim.x)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            S.con_lambda[c * 6 + k] = 0.;
            S.con_sys_res[c * 6 + k] = 0.;
        }
    if (t == 0) {
        S.model_status[m] = 2;
        S.model_iters[m] = 0;
        S.model_active[m] = 1;
    }
}

__global__ void k_plain_clear_flag(DevSys S) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *S.any_active = 0;
}

// node -> model through the block table; nodes without dofs take no part in the iteration
__device__ __forceinline__ int node_model(const DevSys& S, int node) {
    const int b = S.node_block[node];
    return b < 0 ? -1 : S.block_model[b];
}

// UpdateTangentOperator (src/step/update_tangent_operator.hpp:13-26)
__global__ void k_plain_tangent(DevSys S) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= S.num_nodes) return;
    const int m = node_model(S, node);
    if (m < 0 || !S.model_active[m]) return;
    double qd[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) qd[c] = S.q_delta[node * 6 + c];
    tangent_rot_block(S.p.h, qd, S.tan_rot + node * 9);
}

// UpdateSystemVariablesBeams (src/step/update_system_variables_beams.hpp:13-69)
__global__ void __launch_bounds__(kThreads) k_plain_element(DevSys S, int Nm, int Qm) {
    extern __shared__ __align__(16) char sm[];
    const Lay L = make_layout(Nm, Qm, 1, 0, false, 8, false);
    const int e = blockIdx.x;
    const int m = S.block_model[S.node_block[S.elem_nodes[S.elem_node_off[e]]]];
    if (!S.model_active[m]) return;
    element_phase<double, false>(S, e, sm, L, 0, S.elem_res + (size_t)S.elem_node_off[e] * 6,
                                 S.elem_mat + S.elem_mat_off[e]);
}

// FP32-tangent option of the plain form: QP matrices in FP32 and the node-pair contraction on
// the matrix cores (as in the fused FP32 form), element terms stored in FP64.
__global__ void __launch_bounds__(kThreads) k_plain_element_f32(DevSys S, int Nm, int Qm) {
    extern __shared__ __align__(16) char sm[];
    const Lay L = make_layout(Nm, Qm, 1, 0, false, 4, false);
    const int e = blockIdx.x;
    const int m = S.block_model[S.node_block[S.elem_nodes[S.elem_node_off[e]]]];
    if (!S.model_active[m]) return;
    element_phase<float, false>(S, e, sm, L, 0, S.elem_res + (size_t)S.elem_node_off[e] * 6,
                                S.elem_mat + S.elem_mat_off[e]);
}

// One element's residual and iteration matrix at the current state, with its own tangent
// blocks (the no-allocation element-terms entry point): res [N][6], mat [N][N][36].
__global__ void __launch_bounds__(kThreads) k_element_terms(DevSys S, int Nm, int Qm, int e, double* res, double* mat) {
    extern __shared__ __align__(16) char sm[];
    const Lay L = make_layout(Nm, Qm, 1, 0, false, 8, false);
    element_phase<double, false, true>(S, e, sm, L, 0, res, mat);
}

// UpdateSystemVariablesMasses / UpdateSystemVariablesSprings
// (src/step/update_system_variables_masses.hpp, update_system_variables_springs.hpp):
// one thread per point element, masses first
__global__ void k_plain_points(DevSys S) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S.num_masses) {
        const int node = S.mass_node[k];
        const int m = node_model(S, node);
        if (m < 0 || !S.model_active[m]) return;
        double r[4], ud[3], w[3], wd[3];
#pragma unroll
        for (int c = 0; c < 4; ++c) r[c] = S.q[node * 7 + 3 + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ud[c] = S.vd[node * 6 + c];
            wd[c] = S.vd[node * 6 + 3 + c];
            w[c] = S.v[node * 6 + 3 + c];
        }
        mass_terms(r, ud, w, wd, S.mass_r0 + k * 4, S.mass_Mstar + k * 36, S.model_gravity + m * 3,
                   S.p.beta_prime, S.p.gamma_prime, S.tan_rot + node * 9, S.mass_res + k * 6, S.mass_mat + k * 36);
        return;
    }
    const int j = k - S.num_masses;
    if (j >= S.num_springs) return;
    const int n1 = S.spring_n1[j], n2 = S.spring_n2[j];
    const int m = node_model(S, n1);
    if (m < 0 || !S.model_active[m]) return;
    spring_terms(S.spring_x0 + j * 3, S.q + n1 * 7, S.q + n2 * 7, S.spring_lref[j], S.spring_k[j],
                 S.spring_res + j * 6, S.spring_mat + j * 9);
}

// UpdateConstraintVariables (src/step/update_constraint_variables.hpp:13-37)
__global__ void k_plain_constraints(DevSys S) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= S.num_cons) return;
    const int tn = S.con_target[c], bn = S.con_base[c];
    const int m = node_model(S, tn);
    if (m < 0 || !S.model_active[m]) return;
    const int type = S.con_type[c];
    const bool two = con_two_node(type);
    double input[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) input[k] = S.con_input[c * 7 + k];
    if (con_controlled(type)) input[0] = S.con_control[c];
    double tq[7], bq[7] = {0., 0., 0., 1., 0., 0., 0.};
#pragma unroll
    for (int k = 0; k < 7; ++k) tq[k] = S.q[tn * 7 + k];
    if (two)
#pragma unroll
        for (int k = 0; k < 7; ++k) bq[k] = S.q[bn * 7 + k];
    double res[6], bg[36], tg[36], sres[6] = {0., 0., 0., 0., 0., 0.};
    constraint_eval(type, S.con_X0 + c * 3, S.con_axes + c * 9, input, bq, tq, res, bg, tg, sres);
    const double* lam = S.con_lambda + c * 6;
    double gt[36];
    times_tangent(tg, S.tan_rot + tn * 9, gt);
#pragma unroll
    for (int i = 0; i < 36; ++i) {
        S.con_tgrad[c * 36 + i] = gt[i];
        S.con_tgradT[c * 36 + i] = tg[(i % 6) * 6 + i / 6];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double acc = 0.;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc += tg[k * 6 + i] * lam[k];
        S.con_tlres[c * 6 + i] = acc;
    }
    if (two) {
        times_tangent(bg, S.tan_rot + bn * 9, gt);
#pragma unroll
        for (int i = 0; i < 36; ++i) {
            S.con_bgrad[c * 36 + i] = gt[i];
            S.con_bgradT[c * 36 + i] = bg[(i % 6) * 6 + i / 6];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double acc = 0.;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += bg[k * 6 + i] * lam[k];
            S.con_blres[c * 6 + i] = acc;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        S.con_res[c * 6 + k] = res[k];
        if (type == 4) S.con_sys_res[c * 6 + k] = sres[k];
    }
}

// AssembleSystemResidual / AssembleSystemMatrix / AssembleConstraints{Matrix,Residual}
// (src/step/assemble_*.hpp): dense block, deterministic (elements in order, no atomics)
__global__ void __launch_bounds__(kThreads) k_plain_assemble(DevSys S) {
    __shared__ double rhs[512];
    const int b = blockIdx.x, t = threadIdx.x;
    const int m = S.block_model[b];
    if (!S.model_active[m] || S.block_large[b]) return;
    const int n = S.block_n[b];
    double* A = reinterpret_cast<double*>(S.A) + S.block_A_off[b];
    const int lda = n;
    const double cond = S.p.conditioner;
    for (int k = t; k < n * n; k += blockDim.x) A[k] = 0.;
    for (int i = t; i < n; i += blockDim.x) rhs[i] = 0.;
    __syncthreads();
    // ContributeForcesToVector: b -= f
    for (int k = S.block_node_begin[b] + t; k < S.block_node_end[b]; k += blockDim.x) {
        const int node = S.block_nodes[k];
        const int l = S.node_local[node];
#pragma unroll
        for (int c = 0; c < 6; ++c) rhs[l + c] -= S.f[node * 6 + c];
    }
    __syncthreads();
    // ContributeBeamsToVector / ContributeBeamsToSparseMatrix, elements in order
    for (int ke = S.block_elem_begin[b]; ke < S.block_elem_end[b]; ++ke) {
        const int e = S.block_elems[ke];
        const int N = S.elem_N[e], noff = S.elem_node_off[e];
        for (int k = t; k < 6 * N; k += blockDim.x) {
            const int i = k / 6, c = k % 6;
            rhs[S.node_local[S.elem_nodes[noff + i]] + c] += S.elem_res[(size_t)(noff + i) * 6 + c];
        }
        const double* Me = S.elem_mat + S.elem_mat_off[e];
        for (int k = t; k < N * N * 36; k += blockDim.x) {
            const int c = k % 36, ij = k / 36;
            const int i = ij / N, j = ij % N;
            const int row = S.node_local[S.elem_nodes[noff + i]] + c / 6;
            const int col = S.node_local[S.elem_nodes[noff + j]] + c % 6;
            A[(size_t)row * lda + col] += Me[k] * cond;
        }
        __syncthreads();
    }
    // ContributeMassesToVector / ContributeMassesToSparseMatrix, masses in order
    for (int km = S.block_mass_begin[b]; km < S.block_mass_end[b]; ++km) {
        const int e = S.block_masses[km];
        const int l = S.node_local[S.mass_node[e]];
        if (t < 6) rhs[l + t] += S.mass_res[e * 6 + t];
        if (t < 36) A[(size_t)(l + t / 6) * lda + l + t % 6] += S.mass_mat[e * 36 + t] * cond;
        __syncthreads();
    }
    // ContributeSpringsToVector / ContributeSpringsToSparseMatrix (translational dofs only)
    for (int ks = S.block_spring_begin[b]; ks < S.block_spring_end[b]; ++ks) {
        const int e = S.block_springs[ks];
        const int l1 = S.node_local[S.spring_n1[e]], l2 = S.node_local[S.spring_n2[e]];
        if (t < 6) rhs[(t < 3 ? l1 : l2) + t % 3] += S.spring_res[e * 6 + t];
        __syncthreads();
        if (t < 36) {
            const int blk = t / 9, rc = t % 9;  // node pair (blk / 2, blk % 2), entry rc
            const int lr = (blk / 2) ? l2 : l1, lc = (blk % 2) ? l2 : l1;
            const double sg = (blk == 0 || blk == 3) ? 1. : -1.;
            A[(size_t)(lr + rc / 3) * lda + lc + rc % 3] += sg * S.spring_mat[e * 9 + rc] * cond;
        }
        __syncthreads();
    }
    // constraints (copy_constraints_to_sparse_matrix.hpp, ..._transpose_...)
    const int c0 = S.block_con_begin[b], c1 = S.block_con_end[b];
    for (int k = c0; k < c1; ++k) {
        const int c = S.block_cons[k];
        const int type = S.con_type[c];
        const bool two = con_two_node(type);
        const int rows = S.con_rows[c], lr = S.con_local[c];
        const int lt = S.node_local[S.con_target[c]];
        const int lb = two ? S.node_local[S.con_base[c]] : 0;
        for (int q = t; q < rows * 6; q += blockDim.x) {
            const int i = q / 6, j = q % 6;
            A[(size_t)(lr + i) * lda + lt + j] += S.con_tgrad[c * 36 + i * 6 + j];
            A[(size_t)(lt + j) * lda + lr + i] += S.con_tgradT[c * 36 + j * 6 + i];
            if (two) {
                A[(size_t)(lr + i) * lda + lb + j] += S.con_bgrad[c * 36 + i * 6 + j];
                A[(size_t)(lb + j) * lda + lr + i] += S.con_bgradT[c * 36 + j * 6 + i];
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        for (int k = c0; k < c1; ++k) {
            const int c = S.block_cons[k];
            const int lt = S.node_local[S.con_target[c]];
            for (int j = 0; j < 6; ++j) rhs[lt + j] += S.con_sys_res[c * 6 + j];
        }
        for (int k = c0; k < c1; ++k) {
            const int c = S.block_cons[k];
            const int type = S.con_type[c];
            if (con_two_node(type)) {
                const int lb = S.node_local[S.con_base[c]];
                for (int j = 0; j < 6; ++j) rhs[lb + j] += S.con_blres[c * 6 + j];
            }
            const int lt = S.node_local[S.con_target[c]];
            for (int j = 0; j < 6; ++j) rhs[lt + j] += S.con_tlres[c * 6 + j];
        }
        for (int k = c0; k < c1; ++k) {
            const int c = S.block_cons[k];
            const int lr = S.con_local[c];
            for (int i = 0; i < S.con_rows[c]; ++i) rhs[lr + i] += S.con_res[c * 6 + i];
        }
    }
    __syncthreads();
    pin_rotations(S, b, A, lda, rhs);
    condition_rhs(S, b, rhs);
}

// SolveSystem (src/step/solve_system.hpp:14-39): LU + solve per block in LDS
__global__ void __launch_bounds__(kThreads) k_plain_solve(DevSys S, int nm) {
    extern __shared__ __align__(16) char sm[];
    const Lay L = make_layout(1, 1, nm, 0, false, 8, true);
    const int b = blockIdx.x, t = threadIdx.x;
    const int m = S.block_model[b];
    if (!S.model_active[m] || S.block_large[b]) return;
    const int n = S.block_n[b];
    double* A = reinterpret_cast<double*>(sm + L.A);
    double* rhs = reinterpret_cast<double*>(sm + L.rhs);
    double* xs = reinterpret_cast<double*>(sm + L.xs);
    double* xs2 = reinterpret_cast<double*>(sm + L.xs2);
    int* piv = reinterpret_cast<int*>(sm + L.piv);
    const double* Ag = reinterpret_cast<const double*>(S.A) + S.block_A_off[b];
    for (int k = t; k < n * n; k += blockDim.x) A[(k / n) * L.lda + (k % n)] = Ag[k];
    for (int i = t; i < n; i += blockDim.x) rhs[i] = S.rhs[S.block_vec_off[b] + i];
    __syncthreads();
    factor_solve_uncondition<double>(S, b, A, nullptr, L.lda, piv, rhs, xs, xs2, sm, L);
}

// Large blocks: position of entry (row, col) among the block's CRS nonzeros (columns ascending
// in every row: a binary search), -1 if the graph does not hold it.
__device__ __forceinline__ int64_t crs_find(const DevSys& S, int b, int row, int col) {
    const int* rp = S.crs_rp + S.crs_rp_off[b];
    const int* ci = S.crs_col + S.crs_nz_off[b];
    int lo = rp[row], hi = rp[row + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && ci[lo] == col) ? S.crs_nz_off[b] + lo : -1;
}

// Every nonzero of the large blocks to zero (before each assembly)
__global__ void k_crs_zero(DevSys S) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S.crs_nnz; i += (int64_t)gridDim.x * blockDim.x)
        S.crs_val[i] = 0.;
}

// Large blocks: assembly into the CRS values, one workgroup per item.  A node item owns its 6
// columns and its 6 rhs entries, a constraint item its rows' columns and rhs entries; each
// entry receives its contributions in the order of k_plain_assemble (elements in order, masses,
// springs, then constraints), so the matrix equals the small-block one.
__global__ void __launch_bounds__(kThreads) k_large_assemble(DevSys S) {
    const int item = S.big_item[blockIdx.x], t = threadIdx.x;
    const bool is_node = item >= 0;
    const int node = is_node ? item : -1, con = is_node ? -1 : ~item;
    const int b = is_node ? S.node_block[node] : S.node_block[S.con_target[con]];
    const int m = S.block_model[b];
    if (!S.model_active[m]) return;
    const int nsys = S.block_nsys[b], off = S.block_vec_off[b];
    double* V = S.crs_val;
    const double cond = S.p.conditioner;
    // entry (row, col) += v; every (row, col) an item adds to lies in the graph
    auto add = [&](int row, int col, double v) {
        const int64_t p = crs_find(S, b, row, col);
        if (p >= 0) V[p] += v;
    };
    if (!is_node) {
        const int rows = S.con_rows[con], lr = S.con_local[con];
        const int type = S.con_type[con];
        const bool two = con_two_node(type);
        const int lt = S.node_local[S.con_target[con]];
        const int lb = two ? S.node_local[S.con_base[con]] : 0;
        // B^T: the constraint's columns, rows of its target (and base) dofs
        for (int q = t; q < rows * 6; q += blockDim.x) {
            const int i = q / 6, j = q % 6;
            add(lt + j, lr + i, S.con_tgradT[con * 36 + j * 6 + i]);
            if (two) add(lb + j, lr + i, S.con_bgradT[con * 36 + j * 6 + i]);
        }
        if (t < rows) {
            const double raw = 0. + S.con_res[con * 6 + t];
            S.rhs_raw[off + lr + t] = raw;
            S.rhs[off + lr + t] = -1. * raw;  // constraint rows are not scaled
        }
        return;
    }
    const int l = S.node_local[node];
    // beams (ContributeBeamsToSparseMatrix), elements in order
    for (int q = S.node_elem_begin[node]; q < S.node_elem_begin[node + 1]; ++q) {
        const int e = S.node_elem[2 * q], jl = S.node_elem[2 * q + 1];
        const int N = S.elem_N[e], noff = S.elem_node_off[e];
        const double* Me = S.elem_mat + S.elem_mat_off[e];
        for (int k = t; k < N * 36; k += blockDim.x) {
            const int i = k / 36, rc = k % 36;
            const int row = S.node_local[S.elem_nodes[noff + i]] + rc / 6;
            add(row, l + rc % 6, Me[(size_t)(i * N + jl) * 36 + rc] * cond);
        }
        __syncthreads();
    }
    // masses (diagonal node block)
    for (int q = S.node_mass_begin[node]; q < S.node_mass_begin[node + 1]; ++q) {
        const int e = S.node_mass[q];
        if (t < 36) add(l + t / 6, l + t % 6, S.mass_mat[e * 36 + t] * cond);
        __syncthreads();
    }
    // springs: this node's column block of [a, -a; -a, a]
    for (int q = S.node_spring_begin[node]; q < S.node_spring_begin[node + 1]; ++q) {
        const int e = S.node_spring[q];
        const int n1 = S.spring_n1[e], n2 = S.spring_n2[e];
        const int l1 = S.node_local[n1], l2 = S.node_local[n2];
        if (t < 18) {
            const int half = t / 9, rc = t % 9;  // row node (n1, n2) of the column node
            const int lr = half ? l2 : l1;
            const bool diag = (half == 0) == (node == n1);
            add(lr + rc / 3, l + rc % 3, (diag ? 1. : -1.) * S.spring_mat[e * 9 + rc] * cond);
        }
        __syncthreads();
    }
    // constraint gradients in this node's columns (B T: constraint rows, node dofs)
    for (int q = S.node_con_begin[node]; q < S.node_con_begin[node + 1]; ++q) {
        const int c = S.node_con[q];
        const int rows = S.con_rows[c], lr = S.con_local[c];
        const bool tgt = S.con_target[c] == node;
        const double* G = tgt ? S.con_tgrad + c * 36 : S.con_bgrad + c * 36;
        for (int k = t; k < rows * 6; k += blockDim.x) {
            const int i = k / 6, j = k % 6;
            add(lr + i, l + j, G[i * 6 + j]);
        }
        __syncthreads();
    }
    const bool pin = S.node_pin[node] != 0;  // translations only: rotational rows held at zero
    if (pin && t >= 3 && t < 6) {
        const int64_t p = crs_find(S, b, l + t, l + t);
        if (p >= 0) V[p] = 1.;
    }
    // residual entries of this node
    if (t < 6) {
        double r = 0.;
        r -= S.f[node * 6 + t];
        for (int q = S.node_elem_begin[node]; q < S.node_elem_begin[node + 1]; ++q) {
            const int e = S.node_elem[2 * q], jl = S.node_elem[2 * q + 1];
            r += S.elem_res[(size_t)(S.elem_node_off[e] + jl) * 6 + t];
        }
        for (int q = S.node_mass_begin[node]; q < S.node_mass_begin[node + 1]; ++q) r += S.mass_res[S.node_mass[q] * 6 + t];
        for (int q = S.node_spring_begin[node]; q < S.node_spring_begin[node + 1]; ++q) {
            const int e = S.node_spring[q];
            if (t < 3) r += S.spring_res[e * 6 + (S.spring_n1[e] == node ? 0 : 3) + t];
        }
        for (int q = S.node_con_begin[node]; q < S.node_con_begin[node + 1]; ++q) {
            const int c = S.node_con[q];
            if (S.con_target[c] == node) r += S.con_sys_res[c * 6 + t];
        }
        for (int q = S.node_con_begin[node]; q < S.node_con_begin[node + 1]; ++q) {
            const int c = S.node_con[q];
            const int type = S.con_type[c];
            if (con_two_node(type) && S.con_base[c] == node) r += S.con_blres[c * 6 + t];
            if (S.con_target[c] == node) r += S.con_tlres[c * 6 + t];
        }
        if (pin && t >= 3) r = 0.;
        S.rhs_raw[off + l + t] = r;
        S.rhs[off + l + t] = -1. * (l + t < nsys ? r * cond : r);
    }
}

// Condensed solve of a large block (static condensation onto the interface z: free nodes and
// constraint rows).  Element e's dofs couple to the rest only through its constraint rows, so
// with K_e its diagonal block, B_e the constraint rows touching it and b_e its rhs:
//   x_e = K_e^-1 (b_e - B_e^T lambda),  (Z - sum_e B_e K_e^-1 B_e^T) z = r_z - sum_e B_e K_e^-1 b_e.
// k_cond_factor: one workgroup per element, LU of [K_e | B_e^T | b_e] in LDS (the blocked LU
// above, several right-hand sides), Y = K_e^-1 [B_e^T | b_e] -> Ybuf, S_e = B_e Y -> Sbuf.

__host__ __device__ inline int cond_lds_bytes(int ne, int me) {
    const int lda = (ne + me + 1) | 1;
    return align16(ne * lda * 8) + 2 * kLuTile * 4 + 2 * 256 * 8 + 16 + align16(ne * 4) + ne * (me + 1) * 8;
}

// rin: right-hand side vector in the layout of S.rhs (the conditioned rhs, or a refinement
// residual).  TT: storage type of the element factors (FP64, or FP32 with FP64 accumulation in
// the triangular solves).
template <typename TT>
__device__ __forceinline__ void cond_factor_body(DevSys& S, const double* rin) {
    extern __shared__ __align__(16) char sm[];
    const int ce = blockIdx.x, t = threadIdx.x;
    const int b = S.cond_elem_block[ce];
    if (!S.model_active[S.block_model[b]]) return;
    const int vb = S.block_vec_off[b];
    const double* V = S.crs_val + S.crs_nz_off[b];
    const int r0 = S.cond_rows_off[ce], ne = S.cond_rows_off[ce + 1] - r0;
    const int l0 = S.cond_lam_off[ce], me = S.cond_lam_off[ce + 1] - l0;
    const int* rows = S.cond_rows + r0;
    const int nrhs = me + 1;
    const int lda = (ne + nrhs) | 1;
    TT* M = reinterpret_cast<TT*>(sm);
    int off = align16(ne * lda * 8);
    int* ph = reinterpret_cast<int*>(sm + off);  // two buffers
    off += 2 * kLuTile * 4;
    void* lc = sm + off;
    off += 2 * 256 * 8;
    int* sync = reinterpret_cast<int*>(sm + off);
    off += 16;
    int* piv = reinterpret_cast<int*>(sm + off);
    off += align16(ne * 4);
    double* Yl = reinterpret_cast<double*>(sm + off);  // LDS copy of Y for S_e
    const double* rhs = rin + vb;
    for (int k = t; k < ne * lda; k += blockDim.x) M[k] = TT(0);
    __syncthreads();
    // the element rows' CRS entries into [K_e | B_e^T] through the static gather map (thread per
    // entry, four independent loads in flight per thread); the rhs column last
    {
        const int g0 = S.cond_gat_off[ce], g1 = S.cond_gat_off[ce + 1];
        const int* src = S.cond_gat_src;
        const int* dst = S.cond_gat_dst;
        for (int q = g0 + t; q < g1; q += 8 * blockDim.x) {
            int sq[8], dq[8];
            double vq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int qq = q + u * (int)blockDim.x;
                sq[u] = qq < g1 ? src[qq] : src[q];
                dq[u] = qq < g1 ? dst[qq] : -1;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) vq[u] = V[sq[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (dq[u] >= 0) M[dq[u]] = (TT)vq[u];
        }
        for (int i = t; i < ne; i += blockDim.x) M[i * lda + ne + me] = (TT)rhs[rows[i]];
    }
    __syncthreads();
    const LuScratch W{ph, lc, sync, S.lu_wait_limit, S.dev_error};
    lu_factor_augmented<TT>(M, ne, lda, piv, W, nrhs);
    double* Y = S.Ybuf + S.cond_Y_off[ce];
    back_substitute_cols_blocked<TT>(M, ne, lda, ph, nrhs, Yl, Y, ne);
    __syncthreads();
    double* Sm = S.Sbuf + S.cond_S_off[ce];
    for (int q = t; q < me * nrhs; q += blockDim.x) {  // S_e = B_e Y: constraint row lam[k]'s entries in this element
        const int k = q % me, c = q / me;
        const int u0 = S.cond_bt_off[l0 + k], u1 = S.cond_bt_off[l0 + k + 1];
        double acc = 0.;
        for (int u = u0; u < u1; u += 4) {  // four entries' loads in flight, summed in order
            int sq[4], cq[4];
            double vq[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int uu = u + w < u1 ? u + w : u;
                sq[w] = S.cond_bt_src[uu];
                cq[w] = S.cond_bt_col[uu];
            }
#pragma unroll
            for (int w = 0; w < 4; ++w) vq[w] = V[sq[w]];
#pragma unroll
            for (int w = 0; w < 4; ++w)
                if (u + w < u1) acc += vq[w] * Yl[(size_t)c * ne + cq[w]];
        }
        Sm[(size_t)c * me + k] = acc;
    }
}
__global__ void __launch_bounds__(kThreads) k_cond_factor_f64(DevSys S, const double* rin) {
    cond_factor_body<double>(S, rin);
}
__global__ void __launch_bounds__(kThreads) k_cond_factor_f32(DevSys S, const double* rin) {
    cond_factor_body<float>(S, rin);
}

// Interface system of block b: Z = A[z, z] - sum_e S_e, r_z - sum_e t_e, in three launches:
//  k_cond_interface_clear  (grid-stride) the entries the assembly and the solve touch: for a
//                          tree-structured interface its vertices' diagonal and parent-child
//                          blocks, else the whole dense matrix;
//  k_cond_interface_fill   one wave per interface row: the row's CRS entries in interface columns,
//                          and its rhs entry (rows are disjoint, so no ordering);
//  k_cond_interface        the elements' S_e and t_e, one launch per colour (elements of one
//                          colour share no constraint row; every entry receives its updates in
//                          colour order), thread per entry.
__global__ void k_cond_interface_clear(DevSys S, int b) {
    if (!S.model_active[S.block_model[b]]) return;
    const int z0 = S.block_z_off[b], nz = S.block_z_off[b + 1] - z0;
    double* Z = S.Zbuf + S.block_Z_off[b];
    const int nv = S.tree_block[6 * b + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t q0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nv > 0) {
        const int v0 = S.tree_block[6 * b];
        for (int64_t q = q0; q < (int64_t)nv * kTreeMaxBlock * kTreeMaxBlock; q += stride) {
            const int v = v0 + (int)(q / (kTreeMaxBlock * kTreeMaxBlock)), e = (int)(q % (kTreeMaxBlock * kTreeMaxBlock));
            const int i = e / kTreeMaxBlock, j = e % kTreeMaxBlock;
            const int s = S.tree_vsize[v], p = S.tree_vparent[v];
            const int* rv = S.tree_vrows + S.tree_vrow_off[v];
            if (i < s && j < s) Z[(size_t)rv[j] * nz + rv[i]] = 0.;
            if (p >= 0 && i < s && j < S.tree_vsize[p]) {
                const int* rp = S.tree_vrows + S.tree_vrow_off[p];
                Z[(size_t)rp[j] * nz + rv[i]] = 0.;
                Z[(size_t)rv[i] * nz + rp[j]] = 0.;
            }
        }
    } else {
        for (int64_t q = q0; q < (int64_t)nz * nz; q += stride) Z[q] = 0.;
    }
}

__global__ void k_cond_interface_fill(DevSys S, int b, const double* rin) {
    if (!S.model_active[S.block_model[b]]) return;
    const int a = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    const int z0 = S.block_z_off[b], nz = S.block_z_off[b + 1] - z0, vb = S.block_vec_off[b];
    if (a >= nz) return;
    const int row = S.cond_z[z0 + a];
    const int* rp = S.crs_rp + S.crs_rp_off[b];
    const int* ci = S.crs_col + S.crs_nz_off[b];
    const double* V = S.crs_val + S.crs_nz_off[b];
    double* Z = S.Zbuf + S.block_Z_off[b];
    for (int p = rp[row] + lane; p < rp[row + 1]; p += 64) {
        const int zc = S.dof_zidx[vb + ci[p]];
        if (zc >= 0) Z[(size_t)zc * nz + a] = V[p];
    }
    if (lane == 0) S.zrhs[z0 + a] = rin[vb + row];
}

__global__ void k_cond_interface(DevSys S, int b, int col) {
    if (!S.model_active[S.block_model[b]]) return;
    const int z0 = S.block_z_off[b], nz = S.block_z_off[b + 1] - z0;
    double* Z = S.Zbuf + S.block_Z_off[b];
    double* rz = S.zrhs + z0;
    constexpr int kStride = kCondMaxLam * (kCondMaxLam + 1);
    const int c0 = S.cond_color_off[S.block_color[2 * b] + col], c1 = S.cond_color_off[S.block_color[2 * b] + col + 1];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;  // entry of one element of this colour
    if (q >= (c1 - c0) * kStride) return;
    const int ce = S.cond_color_ce[c0 + q / kStride], e = q % kStride;
    const int l0 = S.cond_lam_off[ce], me = S.cond_lam_off[ce + 1] - l0;
    const int* lz = S.cond_lam_z + l0;
    const double* Sm = S.Sbuf + S.cond_S_off[ce];
    if (e < me * me) Z[(size_t)lz[e / me] * nz + lz[e % me]] -= Sm[e];
    else if (e < me * me + me) rz[lz[e - me * me]] -= Sm[(size_t)me * me + (e - me * me)];
}

// r = b - A x over one large block's rows (CRS): one wave per row, lanes over the row's
// entries, a wave reduction (fixed shape, so the sum is deterministic)
__global__ void k_crs_residual(DevSys S, int b, const double* x, double* r) {
    const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (row >= S.block_n[b] || !S.model_active[S.block_model[b]]) return;
    const int vb = S.block_vec_off[b];
    const int* rp = S.crs_rp + S.crs_rp_off[b];
    const int* ci = S.crs_col + S.crs_nz_off[b];
    const double* V = S.crs_val + S.crs_nz_off[b];
    double acc = 0.;
    for (int p = rp[row] + lane; p < rp[row + 1]; p += 64) acc = fma(V[p], x[vb + ci[p]], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) r[vb + row] = S.rhs[vb + row] - acc;
}

// A block's CRS values into its dense column-major matrix (zeroed beforehand): the library-LU
// path of blocks that do not condense
__global__ void k_crs_to_dense(DevSys S, int b) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.block_n[b];
    if (row >= n || !S.model_active[S.block_model[b]]) return;
    const int* rp = S.crs_rp + S.crs_rp_off[b];
    const int* ci = S.crs_col + S.crs_nz_off[b];
    const double* V = S.crs_val + S.crs_nz_off[b];
    double* A = S.Abig + S.block_big_off[b];
    for (int p = rp[row]; p < rp[row + 1]; ++p) A[(size_t)ci[p] * n + row] = V[p];
}

// Tree-structured interface system of a condensed block (see the host's tree schedule): block
// elimination from the leaves, one thread per vertex of a level (vertices of one height are
// independent), then substitution from the roots.  Per vertex v with s <= kTreeMaxBlock rows:
//   S_v = Z_vv - sum_children Z_vc X_c,  [X_v | zhat_v] = S_v^-1 [Z_vp | r_v - sum_children Z_vc zhat_c]
// (LU with partial pivoting of S_v: largest |a|, lowest row on ties, multipliers and the back
// substitution scaled by the pivot's reciprocal; rows past s padded with the identity), then
// z_v = zhat_v - X_v z_p.  Z is the dense column-major interface matrix of
// k_cond_interface, zrhs its right-hand side, overwritten by z.  Every vertex's three blocks
// (Z_vv, Z_v,parent, Z_parent,v) and rhs are first copied to LDS in one parallel pass; X, zhat
// and z live in LDS too.
constexpr int kTreeThreads = 256;

__global__ void __launch_bounds__(kTreeThreads) k_tree_interface(DevSys S, int b) {
    extern __shared__ __align__(16) char sm[];
    if (!S.model_active[S.block_model[b]]) return;
    constexpr int K = kTreeMaxBlock, KK = kTreeMaxBlock * kTreeMaxBlock;
    const int* tb = S.tree_block + 6 * b;
    const int vb = tb[0], nv = tb[1];
    const int z0 = S.block_z_off[b], nz = S.block_z_off[b + 1] - z0;
    const double* Z = S.Zbuf + S.block_Z_off[b];
    double* r = S.zrhs + z0;
    double* Ds = reinterpret_cast<double*>(sm);  // [nv][K][K] Z_vv
    double* Up = Ds + KK * nv;                   // [nv][K][K] Z_v,parent
    double* Lp = Up + KK * nv;                   // [nv][K][K] Z_parent,v
    double* Xs = Lp + KK * nv;                   // [nv][K][K] X_v
    double* xh = Xs + KK * nv;                   // [nv][K]: the rhs, then zhat_v
    double* zs = xh + K * nv;                    // [nv][K]: z_v
    // the block's schedule in LDS as 16-bit local ids (vertex sizes, parents, child lists, rows,
    // elimination and substitution orders with their level offsets), read once from HBM
    const int nle = tb[3], nlb = tb[5];
    short* m_sz = reinterpret_cast<short*>(zs + K * nv);
    short* m_par = m_sz + nv;
    short* m_coff = m_par + nv;      // [nv + 1]
    short* m_ch = m_coff + nv + 1;   // [nv]
    short* m_row = m_ch + nv;        // [nv][K], zero past the vertex's size
    short* m_el = m_row + nv * K;    // [nv]
    short* m_ellv = m_el + nv;       // [nle + 1]
    short* m_bk = m_ellv + nle + 1;  // [nv]
    short* m_bklv = m_bk + nv;       // [nlb + 1]
    {
        const int c0 = S.tree_vchild_off[vb], nch = S.tree_vchild_off[vb + nv] - c0;
        const int e0 = S.tree_elim_lvl[tb[2]], k0 = S.tree_back_lvl[tb[4]];
        for (int q = threadIdx.x; q < nv; q += blockDim.x) {
            const int v = vb + q, sz = S.tree_vsize[v], par = S.tree_vparent[v];
            const int* rows = S.tree_vrows + S.tree_vrow_off[v];
            m_sz[q] = (short)sz;
            m_par[q] = (short)(par >= 0 ? par - vb : -1);
            m_coff[q] = (short)(S.tree_vchild_off[v] - c0);
#pragma unroll
            for (int i = 0; i < K; ++i) m_row[q * K + i] = (short)(i < sz ? rows[i] : 0);
            m_el[q] = (short)(S.tree_elim[e0 + q] - vb);
            m_bk[q] = (short)(S.tree_back[k0 + q] - vb);
        }
        for (int q = threadIdx.x; q < nch; q += blockDim.x) m_ch[q] = (short)(S.tree_vchildren[c0 + q] - vb);
        for (int q = threadIdx.x; q <= nle; q += blockDim.x) m_ellv[q] = (short)(S.tree_elim_lvl[tb[2] + q] - e0);
        for (int q = threadIdx.x; q <= nlb; q += blockDim.x) m_bklv[q] = (short)(S.tree_back_lvl[tb[4] + q] - k0);
        if (threadIdx.x == 0) m_coff[nv] = (short)nch;
    }
    __syncthreads();
    // ---- every vertex's blocks and rhs from Z into LDS: thread per vertex, each block's 36
    // entries as one batch of loads (indices clamped, values selected after one wait)
    for (int lv = threadIdx.x; lv < nv; lv += blockDim.x) {
        const int s = m_sz[lv], p = m_par[lv];
        const int sp = p >= 0 ? m_sz[p] : 0;
        int rv[K], rp[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            rv[i] = m_row[lv * K + i];
            rp[i] = p >= 0 ? m_row[p * K + i] : 0;
        }
#pragma unroll
        for (int which = 0; which < 3; ++which) {
            const int* ra = which == 2 ? rp : rv;  // block rows
            const int* ca = which == 1 ? rp : rv;  // block columns
            const int na = which == 2 ? sp : s, nc = which == 1 ? sp : s;
            double val[KK];
#pragma unroll
            for (int i = 0; i < K; ++i)
#pragma unroll
                for (int j = 0; j < K; ++j) val[i * K + j] = Z[(size_t)ca[j] * nz + ra[i]];
#pragma unroll
            for (int e = 0; e < KK; ++e) pin(val[e]);
            double* dst = (which == 0 ? Ds : which == 1 ? Up : Lp) + lv * KK;
#pragma unroll
            for (int i = 0; i < K; ++i)
#pragma unroll
                for (int j = 0; j < K; ++j) dst[i * K + j] = (i < na && j < nc) ? val[i * K + j] : 0.;
        }
        double rr[K];
#pragma unroll
        for (int i = 0; i < K; ++i) rr[i] = r[rv[i]];
#pragma unroll
        for (int i = 0; i < K; ++i) xh[lv * K + i] = i < s ? rr[i] : 0.;
    }
    __syncthreads();
    // ---- elimination, leaves first: one wave per vertex, lane i < K holding row i of [M | R]
    // (R: the K columns of Z_v,parent, then the rhs); every entry receives the operations of the
    // row-by-row elimination in the same order, lanes >= K compute on a copy of row 0 and store
    // nothing
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63, nwv = (int)blockDim.x >> 6;
    const int li = lane < K ? lane : 0;
    for (int l = 0; l < nle; ++l) {
        const int lo = m_ellv[l], hi = m_ellv[l + 1];
        for (int q = lo + wv; q < hi; q += nwv) {  // uniform per wave
            const int lv = m_el[q];
            const int s = m_sz[lv];
            double M[K], R[K + 1];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                M[j] = (li < s && j < s) ? Ds[lv * KK + li * K + j] : (li == j ? 1. : 0.);
                R[j] = Up[lv * KK + li * K + j];
            }
            R[K] = xh[lv * K + li];
            for (int cq = m_coff[lv]; cq < m_coff[lv + 1]; ++cq) {
                const int lc = m_ch[cq], sc = m_sz[lc];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (k < sc) {  // uniform
                        const double* Xk = Xs + lc * KK + k * K;  // row k of X_c (one address for all lanes)
                        const double hk = xh[lc * K + k];
                        const double zik = Lp[lc * KK + li * K + k];  // Z_v,c[i][k]: zero past v's rows
#pragma unroll
                        for (int j = 0; j < K; ++j) M[j] = fma(-zik, Xk[j], M[j]);
                        R[K] = fma(-zik, hk, R[K]);
                    }
                }
            }
            // LU with partial pivoting (largest |m_ij| over rows i >= j, the first on ties), the
            // interchanges and eliminations applied to R
#pragma unroll
            for (int j = 0; j < K; ++j) {
                int pr = j;
                double best = fabs(rdlane(M[j], j));
#pragma unroll
                for (int i = j + 1; i < K; ++i) {
                    const double cand = fabs(rdlane(M[j], i));
                    if (cand > best) {
                        best = cand;
                        pr = i;
                    }
                }
                if (pr != j) {  // uniform: rows j and pr trade places
                    const int src = lane == j ? pr : (lane == pr ? j : lane);
#pragma unroll
                    for (int c = 0; c < K; ++c) M[c] = __shfl(M[c], src, 64);
#pragma unroll
                    for (int c = 0; c <= K; ++c) R[c] = __shfl(R[c], src, 64);
                }
                double pj[K], rj[K + 1];
#pragma unroll
                for (int c = 0; c < K; ++c) pj[c] = rdlane(M[c], j);
#pragma unroll
                for (int c = 0; c <= K; ++c) rj[c] = rdlane(R[c], j);
                const double inv = 1. / pj[j];  // one division per pivot; the multipliers and the
                if (lane == j) M[j] = inv;        // back substitution scale by it
                if (lane > j) {
                    const double lij = M[j] * inv;
#pragma unroll
                    for (int c = j + 1; c < K; ++c) M[c] = fma(-lij, pj[c], M[c]);
#pragma unroll
                    for (int c = 0; c <= K; ++c) R[c] = fma(-lij, rj[c], R[c]);
                }
            }
#pragma unroll
            for (int j = K - 1; j >= 0; --j) {
                if (lane == j)
#pragma unroll
                    for (int c = 0; c <= K; ++c) R[c] = R[c] * M[j];  // M[j] holds 1 / u_jj on lane j
                double rj[K + 1];
#pragma unroll
                for (int c = 0; c <= K; ++c) rj[c] = rdlane(R[c], j);
                if (lane < j)
#pragma unroll
                    for (int c = 0; c <= K; ++c) R[c] = fma(-M[j], rj[c], R[c]);
            }
            if (lane < K) {
#pragma unroll
                for (int j = 0; j < K; ++j) Xs[lv * KK + lane * K + j] = R[j];
                xh[lv * K + lane] = R[K];
            }
        }
        __syncthreads();
    }
    // ---- substitution, roots first
    for (int l = 0; l < nlb; ++l) {  // one wave per vertex, lane i < size its row i
        const int lo = m_bklv[l], hi = m_bklv[l + 1];
        for (int q = lo + wv; q < hi; q += nwv) {
            const int lv = m_bk[q];
            const int s = m_sz[lv], p = m_par[lv];
            const int sp = p >= 0 ? m_sz[p] : 0;
            if (lane < s) {
                double zi = xh[lv * K + lane];
#pragma unroll
                for (int j = 0; j < K; ++j)
                    if (j < sp) zi = fma(-Xs[lv * KK + lane * K + j], zs[p * K + j], zi);
                zs[lv * K + lane] = zi;
                r[m_row[lv * K + lane]] = zi;
            }
        }
        __syncthreads();
    }
}

// Element dofs from the interface solution: x_e = Y_b - Y_B lambda_e; interface rows copied
__global__ void __launch_bounds__(kThreads) k_cond_backsub(DevSys S, double* xout) {
    const int ce = blockIdx.x, t = threadIdx.x;
    const int b = S.cond_elem_block[ce];
    if (!S.model_active[S.block_model[b]]) return;
    const int r0 = S.cond_rows_off[ce], ne = S.cond_rows_off[ce + 1] - r0;
    const int l0 = S.cond_lam_off[ce], me = S.cond_lam_off[ce + 1] - l0;
    const int* rows = S.cond_rows + r0;
    const int* lz = S.cond_lam_z + l0;
    const double* Y = S.Ybuf + S.cond_Y_off[ce];
    const double* z = S.zrhs + S.block_z_off[b];
    double* x = xout + S.block_vec_off[b];
    for (int i = t; i < ne; i += blockDim.x) {
        double v = Y[(size_t)me * ne + i];
        for (int k = 0; k < me; ++k) v -= Y[(size_t)k * ne + i] * z[lz[k]];
        x[rows[i]] = v;
    }
}
__global__ void k_cond_scatter_z(DevSys S, int b, double* xout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int z0 = S.block_z_off[b], nz = S.block_z_off[b + 1] - z0;
    if (i >= nz || !S.model_active[S.block_model[b]]) return;
    xout[S.block_vec_off[b] + S.cond_z[z0 + i]] = S.zrhs[z0 + i];
}
// refinement bookkeeping of one block: x += d (x non-null), r = rhs (r non-null, before r -= A x)
__global__ void k_block_axpy_copy(DevSys S, int b, const double* d, double* x, double* r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.block_n[b] || !S.model_active[S.block_model[b]]) return;
    const int k = S.block_vec_off[b] + i;
    if (x) x[k] += d[k];
    if (r) r[k] = S.rhs[k];
}

// UnconditionSolution for a large block: x (in xsol) -> lambda rows divided by the conditioner
__global__ void k_large_uncondition(DevSys S, int b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.block_n[b], nsys = S.block_nsys[b], off = S.block_vec_off[b];
    if (i >= n || !S.model_active[S.block_model[b]]) return;
    if (i >= nsys) S.xsol[off + i] /= S.p.conditioner;
}

// HostState::CopyFromState in one transfer: per node x[7], q[7], v[6], vd[6] -> out[n][26]
__global__ void k_pack_motion(DevSys S, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.num_nodes) return;
    double* o = out + (size_t)i * 26;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        o[c] = S.x[i * 7 + c];
        o[7 + c] = S.q[i * 7 + c];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        o[14 + c] = S.v[i * 6 + c];
        o[20 + c] = S.vd[i * 6 + c];
    }
}

// mixed precision helpers for the large-block refinement
__global__ void k_f64_to_f32(const double* x, float* y, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) y[i] = (float)x[i];
}
__global__ void k_add_f32_to_f64(const float* d, double* x, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) x[i] += (double)d[i];
}

// CalculateConvergenceError + UpdateStatePrediction + UpdateConstraintPrediction
__global__ void __launch_bounds__(kThreads) k_plain_update(DevSys S, int iter) {
    __shared__ double red[16];
    const int m = blockIdx.x, t = threadIdx.x;
    if (!S.model_active[m]) return;
    double err_sum = 0.;
    for (int b = S.model_block_begin[m]; b < S.model_block_end[m]; ++b)
        err_sum += block_error_and_update(S, b, S.xsol + S.block_vec_off[b], red);
    const double err = sqrt(err_sum / (double)S.model_ndofs[m]);
    if (t == 0) {
        S.err_hist[(size_t)m * S.p.max_iter + iter] = err;
        S.model_iters[m] = iter + 1;
        if (!(err > 1.)) {  // step.hpp:54: a NaN error ends the loop as converged
            S.model_status[m] = 0;
            S.model_active[m] = 0;
        } else if (iter + 1 >= S.p.max_iter) {
            S.model_status[m] = 1;
            S.model_active[m] = 0;
        } else {
            *S.any_active = 1;
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_plain_epilogue(DevSys S) {
    const int m = blockIdx.x, t = threadIdx.x;
    if (t == 0) {
        S.model_iter_total[m] += S.model_iters[m];
        S.model_fail_total[m] += S.model_status[m] == 0 ? 0 : 1;
    }
    if (S.model_status[m] != 0) return;
    for (int node = S.model_node_begin[m] + t; node < S.model_node_end[m]; node += blockDim.x)
        epilogue_node(S, node);
    for (int c = S.model_con_begin[m] + t; c < S.model_con_end[m]; c += blockDim.x) constraint_output(S, c);
}

// ---------------------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------------------
size_t fused_lds_bytes(int Nm, int Qm, int nm, int cm, int refine, int precision) {
    return (size_t)make_layout(Nm, Qm, nm, cm, refine > 0, precision ? 4 : 8, true).total;
}
size_t element_lds_bytes(int Nm, int Qm) { return (size_t)make_layout(Nm, Qm, 1, 0, false, 8, false).total; }
size_t solve_lds_bytes(int nm) { return (size_t)make_layout(1, 1, nm, 0, false, 8, true).total; }

hipError_t set_lds_limits(size_t bytes) {
    hipError_t e;
    e = hipFuncSetAttribute((const void*)k_fused_step_f64, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_fused_step_f32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_plain_element, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_element_terms, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_plain_element_f32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_tree_interface, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_cond_factor_f64, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_cond_factor_f32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)k_plain_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

void launch_fused_step(const DevSys& S, int Nm, int Qm, int nm, hipStream_t st) {
    const size_t lds = fused_lds_bytes(Nm, Qm, nm, S.max_block_cons, S.refine_sweeps, S.precision);
    if (S.precision)
        hipLaunchKernelGGL(k_fused_step_f32, dim3(S.num_models), dim3(kThreads), lds, st, S, Nm, Qm, nm);
    else
        hipLaunchKernelGGL(k_fused_step_f64, dim3(S.num_models), dim3(kThreads), lds, st, S, Nm, Qm, nm);
}
// ---------------------------------------------------------------------------------------
// Element-terms scratch of the no-allocation entry point: one element's residual [N][6] and
// iteration matrix [N][N][36] for N <= kMaxElemNodes (128 016 bytes), a file-scope device array,
// one per GPU, held until the process ends.  Host code never passes the array's name to a
// kernel or a copy: it uses the device address looked up once per device below.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kElemTermsScratch];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t element_terms_scratch(int device, double** ptr) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

void launch_element_terms(const DevSys& S, int Nm, int Qm, int e, double* res, double* mat, hipStream_t st) {
    hipLaunchKernelGGL(k_element_terms, dim3(1), dim3(kThreads), element_lds_bytes(Nm, Qm), st, S, Nm, Qm, e, res, mat);
}

void launch_advance_slot(const DevSys& S, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_slot, dim3(1), dim3(64), 0, st, S);
}
void launch_plain_predict(const DevSys& S, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_predict, dim3(S.num_models), dim3(kThreads), 0, st, S);
}
void launch_plain_update(const DevSys& S, int iter, hipStream_t st);
void launch_plain_iteration(const DevSys& S, int Nm, int Qm, int nm, int iter, bool solve,
                            hipStream_t st, hipEvent_t e0, hipEvent_t e1, bool defer_update) {
    hipLaunchKernelGGL(k_plain_clear_flag, dim3(1), dim3(64), 0, st, S);
    hipLaunchKernelGGL(k_plain_tangent, dim3((S.num_nodes + 255) / 256), dim3(256), 0, st, S);
    if (e0) (void)hipEventRecord(e0, st);
    if (S.num_elems > 0) {
        if (S.precision)
            hipLaunchKernelGGL(k_plain_element_f32, dim3(S.num_elems), dim3(kThreads),
                               (size_t)make_layout(Nm, Qm, 1, 0, false, 4, false).total, st, S, Nm, Qm);
        else
            hipLaunchKernelGGL(k_plain_element, dim3(S.num_elems), dim3(kThreads), element_lds_bytes(Nm, Qm), st, S, Nm,
                               Qm);
    }
    if (e1) (void)hipEventRecord(e1, st);
    if (S.num_masses + S.num_springs > 0)
        hipLaunchKernelGGL(k_plain_points, dim3((S.num_masses + S.num_springs + 63) / 64), dim3(64), 0, st, S);
    if (S.num_cons > 0)
        hipLaunchKernelGGL(k_plain_constraints, dim3((S.num_cons + 63) / 64), dim3(64), 0, st, S);
    hipLaunchKernelGGL(k_plain_assemble, dim3(S.num_blocks), dim3(kThreads), 0, st, S);
    if (S.num_big_items > 0) {
        hipLaunchKernelGGL(k_crs_zero, dim3(512), dim3(256), 0, st, S);
        hipLaunchKernelGGL(k_large_assemble, dim3(S.num_big_items), dim3(kThreads), 0, st, S);
    }
    if (!solve) return;
    hipLaunchKernelGGL(k_plain_solve, dim3(S.num_blocks), dim3(kThreads), solve_lds_bytes(nm), st, S, nm);
    if (!defer_update) launch_plain_update(S, iter, st);
}
void launch_plain_update(const DevSys& S, int iter, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_update, dim3(S.num_models), dim3(kThreads), 0, st, S, iter);
}
size_t cond_lds(int ne, int me) { return (size_t)cond_lds_bytes(ne, me); }
void launch_cond_factor(const DevSys& S, size_t lds, bool fp32, const double* rin, hipStream_t st) {
    if (fp32)
        hipLaunchKernelGGL(k_cond_factor_f32, dim3(S.num_cond_elems), dim3(kThreads), lds, st, S, rin);
    else
        hipLaunchKernelGGL(k_cond_factor_f64, dim3(S.num_cond_elems), dim3(kThreads), lds, st, S, rin);
}
void launch_cond_interface(const DevSys& S, int b, int nz, const int* colour_sizes, int ncolour, const double* rin,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_cond_interface_clear, dim3(64), dim3(256), 0, st, S, b);
    hipLaunchKernelGGL(k_cond_interface_fill, dim3((nz + 3) / 4), dim3(256), 0, st, S, b, rin);  // 4 rows per workgroup
    constexpr int kStride = kCondMaxLam * (kCondMaxLam + 1);
    for (int col = 0; col < ncolour; ++col)
        hipLaunchKernelGGL(k_cond_interface, dim3((colour_sizes[col] * kStride + 255) / 256), dim3(256), 0, st, S, b, col);
}
void launch_cond_backsub(const DevSys& S, double* xout, hipStream_t st) {
    hipLaunchKernelGGL(k_cond_backsub, dim3(S.num_cond_elems), dim3(kThreads), 0, st, S, xout);
}
void launch_cond_scatter_z(const DevSys& S, int b, int nz, double* xout, hipStream_t st) {
    hipLaunchKernelGGL(k_cond_scatter_z, dim3((nz + 255) / 256), dim3(256), 0, st, S, b, xout);
}
void launch_block_axpy_copy(const DevSys& S, int b, int n, const double* d, double* x, double* r, hipStream_t st) {
    hipLaunchKernelGGL(k_block_axpy_copy, dim3((n + 255) / 256), dim3(256), 0, st, S, b, d, x, r);
}
// the blocks (4 K x K and 2 K doubles per vertex) and the 16-bit schedule (6 + K entries per
// vertex, one more child offset, and two level-offset lists of at most kTreeMaxVerts + 1)
constexpr size_t tree_lds_bytes() {
    return (size_t)kTreeMaxVerts * (4 * kTreeMaxBlock * kTreeMaxBlock + 2 * kTreeMaxBlock) * sizeof(double) +
           ((size_t)kTreeMaxVerts * (8 + kTreeMaxBlock) + 3) * sizeof(short);
}
static_assert(tree_lds_bytes() <= 160 * 1024, "tree interface solve: LDS");
void launch_tree_interface(const DevSys& S, int b, hipStream_t st) {
    // LDS per vertex of the block's tree: Z_vv, Z_v,parent, Z_parent,v and X (K x K each), the
    // rhs / zhat and z (K each); the host admits a tree only up to kTreeMaxVerts vertices
    hipLaunchKernelGGL(k_tree_interface, dim3(1), dim3(kTreeThreads),
                       tree_lds_bytes(), st,
                       S, b);
}
void launch_crs_residual(const DevSys& S, int b, int n, const double* x, double* r, hipStream_t st) {
    hipLaunchKernelGGL(k_crs_residual, dim3((n + 3) / 4), dim3(256), 0, st, S, b, x, r);  // 4 rows per workgroup
}
void launch_crs_to_dense(const DevSys& S, int b, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_crs_to_dense, dim3((n + 255) / 256), dim3(256), 0, st, S, b);
}
void launch_large_uncondition(const DevSys& S, int b, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_large_uncondition, dim3((n + 255) / 256), dim3(256), 0, st, S, b);
}
void launch_pack_motion(const DevSys& S, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_motion, dim3((S.num_nodes + 255) / 256), dim3(256), 0, st, S, out);
}
void launch_f64_to_f32(const double* x, float* y, size_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, x, y, count);
}
void launch_add_f32_to_f64(const float* d, double* x, int count, hipStream_t st) {
    hipLaunchKernelGGL(k_add_f32_to_f64, dim3((count + 255) / 256), dim3(256), 0, st, d, x, count);
}
void launch_plain_epilogue(const DevSys& S, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_epilogue, dim3(S.num_models), dim3(kThreads), 0, st, S);
}

}  // namespace kyn
