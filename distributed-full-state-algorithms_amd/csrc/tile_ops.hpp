// tile_ops.hpp -- operations on a tile of amplitudes held in LDS, shared by
// the fused-pass kernel (kernels_tile.hip) and the batched circuit kernel
// (kernels_batch.hip).
//
// A tile is 2^m amplitudes whose tile-local index l spells the m tile qubits;
// `base` is the global index of the tile's other bits.  Every function here is
// executed by a whole workgroup of TB threads on the LDS array s and leaves
// LDS consistent only after the caller's barrier.
#pragma once

#include "launch.hpp"

namespace fsa {
namespace tile {


// bit value of a channel qubit: code bit 8 set -> local position in the tile,
// otherwise a global qubit index outside the tile (fixed for the workgroup)
__device__ __forceinline__ uint32_t chan_bit(uint32_t l, uint64_t base, uint32_t code) {
    return (code & 0x100u) ? (l >> (code & 0xffu)) & 1u : (uint32_t)((base >> code) & 1ULL);
}

template <typename R>
__device__ __forceinline__ cx<R> ldc(const double* f, int i) { return mk<R>((R)f[2 * i], (R)f[2 * i + 1]); }

// Amplitudes are stored in precision R and combined in precision C >= R: an
// FP32 tile worked in FP64 is rounded once per op (on its store), and its gate
// factors keep FP64 values (FP32-rounded matrices and angles would perturb
// every amplitude alike, which is the larger error of the two).
template <typename C, typename R>
__device__ __forceinline__ cx<C> up(cx<R> v) { return mk<C>((C)v.x, (C)v.y); }
template <typename R, typename C>
__device__ __forceinline__ cx<R> dn(cx<C> v) { return mk<R>((R)v.x, (R)v.y); }

// s_waitcnt operand for vmcnt(0) with expcnt and lgkmcnt left at their maxima
// (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int WAIT_VMCNT0 = 0x0F70;

// dense k-target gate on LDS.  When the tile holds at least TB groups each
// thread owns whole groups (in place).  Otherwise the 2^K outputs of a group
// are split over TB / groups threads: every thread streams the group's inputs
// from LDS, accumulates its outputs in registers, and writes them after a
// barrier (so no thread overwrites inputs another still reads).  Outputs
// accumulate in input order, as in the streaming kernel.
// tile offset of matrix index l: bit i of l -> tile bit p_i (bits packed 5 per position)
template <int K>
__device__ __forceinline__ uint32_t mat_offset(uint32_t l, uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3,
                                               uint32_t b4) {
    uint32_t o = (l & 1u) ? b0 : 0u;
    if (K > 1 && (l & 2u)) o |= b1;
    if (K > 2 && (l & 4u)) o |= b2;
    if (K > 3 && (l & 8u)) o |= b3;
    if (K > 4 && (l & 16u)) o |= b4;
    return o;
}

template <typename R, int K, int TB, typename C = R>
__device__ __forceinline__ void tile_matk(cx<R>* s, uint32_t T, const TileOp& op, const cx<C>* __restrict__ mat) {
    constexpr int D = 1 << K;
    const uint32_t b0 = 1u << op.p[0];
    const uint32_t b1 = K > 1 ? 1u << op.p[1] : 0u;
    const uint32_t b2 = K > 2 ? 1u << op.p[2] : 0u;
    const uint32_t b3 = K > 3 ? 1u << op.p[3] : 0u;
    const uint32_t b4 = K > 4 ? 1u << op.p[4] : 0u;
    const uint32_t tm = b0 | b1 | b2 | b3 | b4;
#define FSA_OFF(l) mat_offset<K>((l), b0, b1, b2, b3, b4)
    const uint32_t groups = T >> K;
    const uint32_t tid = threadIdx.x;
    constexpr int OUTMAX = sizeof(C) == 8 ? 8 : 16;
    constexpr int OUT = D < OUTMAX ? D : OUTMAX;
    // whole groups per thread, written in place; a full tile has >= TB groups only when D <= OUT
    if constexpr (D <= OUT) if (groups >= (uint32_t)TB) {
        for (uint32_t g = tid; g < groups; g += TB) {
            const cx<C>* m = mat;
            const uint32_t b = ins0_mask32(g, tm);
            cx<C> acc[D];
#pragma unroll
            for (int o = 0; o < D; o++) acc[o] = mk<C>(0, 0);
            // input loop kept rolled: D accumulators and one matrix column live at a time
#pragma unroll 1
            for (int l = 0; l < D; l++) {
                cx<C> x = up<C>(s[b | FSA_OFF((uint32_t)l)]);
#pragma unroll
                for (int o = 0; o < D; o++) acc[o] = cfma(acc[o], m[o * D + l], x);
            }
#pragma unroll
            for (int o = 0; o < D; o++) s[b | FSA_OFF((uint32_t)o)] = dn<R>(acc[o]);
        }
        return;
    }
    // split mode: tpg threads per group, each computing outs = D / tpg <= OUT
    // consecutive outputs; a round covers TB / tpg whole groups, all of whose
    // inputs are read before the barrier and outputs written after it (rounds
    // touch disjoint groups, so they need no barrier between them)
    constexpr uint32_t TPG = D > OUT ? (uint32_t)(D / OUT) : 1u;
    constexpr uint32_t OUTS = (uint32_t)D / TPG;
    constexpr uint32_t GPR = (uint32_t)TB / TPG;             // groups per round
    const uint32_t gl = tid % GPR, sub = tid / GPR;
    const uint32_t r0 = sub * OUTS;
    for (uint32_t g0 = 0; g0 < groups; g0 += GPR) {
        const uint32_t g = g0 + gl;
        const bool active = g < groups;
        const uint32_t b = ins0_mask32(active ? g : 0u, tm);
        cx<C> acc[OUT];
#pragma unroll
        for (int o = 0; o < OUT; o++) acc[o] = mk<C>(0, 0);
        if (GPR >= 64) {
            // a wave spans one `sub`, so its output rows are wave-uniform: read the
            // matrix through the scalar unit instead of one vector load per lane
            const uint32_t r0u = __builtin_amdgcn_readfirstlane(r0);
            const cx<C>* mrow = mat + (size_t)r0u * D;
            if (active) {
#pragma unroll 1
                for (int l = 0; l < D; l++) {
                    cx<C> x = up<C>(s[b | FSA_OFF((uint32_t)l)]);
#pragma unroll
                    for (int o = 0; o < (int)OUTS; o++) acc[o] = cfma(acc[o], mrow[o * D + l], x);
                }
            }
        } else if (active) {
#pragma unroll 1
            for (int l = 0; l < D; l++) {
                cx<C> x = up<C>(s[b | FSA_OFF((uint32_t)l)]);
#pragma unroll
                for (int o = 0; o < (int)OUTS; o++) acc[o] = cfma(acc[o], mat[(r0 + o) * D + l], x);
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int o = 0; o < (int)OUTS; o++) s[b | FSA_OFF(r0 + o)] = dn<R>(acc[o]);
        }
    }
#undef FSA_OFF
}

// ---- dense gates on MFMA ----------------------------------------------------
// Y = U X for the 2^K amplitudes of every group in the tile, as the real GEMM
// [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi] (2D x 2D times 2D x groups, D = 2^K)
// on 16x16x4 MFMA tiles.  Operand layout of v_mfma_*_16x16x4: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element j of lane l is
// column l % 16, row 4 j + l / 16 for f64 and row 4 (l / 16) + j for f32
// (measured on gfx950).  A wave task is one 16-row block
// of the real matrix against four 16-group column blocks; the real matrix (in
// the parameter buffer, row-major 2D x 2D) feeds A, the tile in LDS feeds B.
// All inputs are read before the barrier, outputs written after it.
typedef double fsa_d4 __attribute__((ext_vector_type(4)));
typedef float fsa_f4 __attribute__((ext_vector_type(4)));
template <typename R> struct MfmaAcc;
template <> struct MfmaAcc<double> {
    typedef fsa_d4 T;
    static __device__ __forceinline__ T step(double a, double b, T c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaAcc<float> {
    typedef fsa_f4 T;
    static __device__ __forceinline__ T step(float a, float b, T c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// A tile of T amplitudes holds T / 512 wave tasks for every K (RB * groups / 64),
// so each of the TB / 64 waves takes MAXT = T / (8 TB) of them: 1 for 4096
// amplitudes at TB = 512, 2 for 8192.
//
// The A operand (the real matrix [[Ur, -Ui], [Ui, Ur]]) is read from LDS: M
// holds its left half, the 2D x D reals [Ur; Ui], stored by mat_lds_index.
// LDS reads keep the MFMA loops free of vector-memory loads, whose waits
// would also wait for the next tile's prefetch (loads complete in order).
template <int D>
__device__ __forceinline__ uint32_t mat_lds_index(uint32_t r, uint32_t c) {
    // columns XOR-swizzled by the row: the 16 rows a wave reads at one column
    // fall in distinct banks
    return r * (uint32_t)D + (c ^ (r & (uint32_t)(D - 1)));
}

template <typename R, int K, int TB, int MAXT>
__device__ __forceinline__ void tile_matk_mfma(cx<R>* s, uint32_t T, const TileOp& op, const R* __restrict__ M) {
    typedef typename MfmaAcc<R>::T acc_t;
    constexpr int D = 1 << K, D2 = 2 * D;
    constexpr int RB = D2 / 16;                 // 16-row blocks of the real matrix
    constexpr int WAVES = TB / 64;
    const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1], b2 = 1u << op.p[2];
    const uint32_t b3 = K > 3 ? 1u << op.p[3] : 0u;
    const uint32_t b4 = K > 4 ? 1u << op.p[4] : 0u;
    const uint32_t tm = b0 | b1 | b2 | b3 | b4;
    const uint32_t groups = T >> K;
    const uint32_t cb4 = groups / 64;          // groups of four 16-column blocks
    const uint32_t tasks = RB * cb4;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t lr = lane & 15u, lk = lane >> 4;
    acc_t acc[MAXT][4];
    uint32_t taskId[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        taskId[t] = wave + (uint32_t)t * WAVES;
#pragma unroll
        for (int c = 0; c < 4; c++) acc[t][c] = acc_t{0, 0, 0, 0};
        if (taskId[t] >= tasks) continue;
        const uint32_t rb = taskId[t] % RB, cbase = (taskId[t] / RB) * 4;
        const uint32_t arow = rb * 16 + lr;                // real row of A this lane feeds
        // B-fragment group bases of the four column blocks for this lane
        uint32_t gb[4];
#pragma unroll
        for (int c = 0; c < 4; c++) gb[c] = ins0_mask32((cbase + (uint32_t)c) * 16 + lr, tm);
        const R* sr = reinterpret_cast<const R*>(s);
        auto kstep = [&](int ks, R a) {
            const uint32_t r = (uint32_t)ks * 4 + lk;          // real row of X this lane feeds
            const uint32_t off = mat_offset<K>(r & (D - 1), b0, b1, b2, b3, b4);
            const uint32_t half = r < (uint32_t)D ? 0u : 1u;  // real or imaginary part only
#pragma unroll
            for (int c = 0; c < 4; c++) acc[t][c] = MfmaAcc<R>::step(a, sr[2 * (gb[c] | off) + half], acc[t][c]);
        };
        // not unrolled: four independent 64-cycle FP64 MFMAs per k-step cover
        // the operand latency, and the loop stays clear of register spills.
        // Columns C < D: A[row][C] = left[row][C]; columns C >= D:
        // A[row][C] = -left[row + D][C - D] (row < D, the -Ui block) or
        // left[row - D][C - D] (row >= D, the Ur block)
        const uint32_t r2 = (arow + D) & (D2 - 1);
        const R sgn = arow < (uint32_t)D ? (R)-1 : (R)1;
#pragma unroll 1
        for (int ks = 0; ks < D / 4; ks++) kstep(ks, M[mat_lds_index<D>(arow, (uint32_t)ks * 4 + lk)]);
#pragma unroll 1
        for (int ks = D / 4; ks < D2 / 4; ks++) kstep(ks, sgn * M[mat_lds_index<D>(r2, (uint32_t)ks * 4 + lk - D)]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        if (taskId[t] >= tasks) continue;
        const uint32_t rb = taskId[t] % RB, cbase = (taskId[t] / RB) * 4;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t gbc = ins0_mask32((cbase + (uint32_t)c) * 16 + lr, tm);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // accumulator row j of this lane: FP64 4j + lane/16, FP32 4 (lane/16) + j
                const uint32_t row = rb * 16 + (sizeof(R) == 8 ? 4u * (uint32_t)j + lk : lk * 4u + (uint32_t)j);
                const uint32_t l = row & (D - 1);
                R* dst = reinterpret_cast<R*>(&s[gbc | mat_offset<K>(l, b0, b1, b2, b3, b4)]);
                dst[row < (uint32_t)D ? 0 : 1] = acc[t][c][j];
            }
        }
    }
}

// FP64 gates of 4 or 5 targets in Gauss's three-product form (matrices in LDS,
// 512-thread workgroups: tile_mfma3_form): a wave task is one 16-row block of U
// (D / 16 of them) against two 16-group
// column blocks.  A first sweep over the D / 4 k-steps accumulates P1 = Ur Xr
// and P2 = Ui Xi, which become Yr = P1 - P2 and -(P1 + P2); a second sweep adds
// P3 = (Ur + Ui)(Xr + Xi) to the latter, giving Yi.  Three quarters of the real
// form's matrix operations, the same accumulator registers, and the same LDS
// operand traffic (each sweep reads every amplitude once per row block of U,
// as a 16-byte pair).  The tasks per tile are those of the real form (T / 512
// for every K).  M holds U as complex entries, swizzled like the real form
// (mat_lds_index), so that one 16-byte LDS read gives ur and ui.
// The form is fixed at compile time: with the real form of these gates also
// compiled in, the FP64 kernels exceed 128 VGPRs and spill; the 1024-thread
// kernels spill with it and keep the real form.
template <typename R, int TB>
__host__ __device__ constexpr bool tile_mfma3_form(uint32_t k) {
    return sizeof(R) == 8 && TB == 512 && (k == 4 || k == 5);
}

template <int K, int TB, int MAXT>
__device__ __forceinline__ void tile_matk_mfma3(cx<double>* s, uint32_t T, const TileOp& op,
                                                const cx<double>* __restrict__ M) {
    typedef MfmaAcc<double>::T acc_t;
    static_assert(K == 4 || K == 5, "three-product form: 16 or 32 rows of U");
    constexpr int D = 1 << K;
    constexpr int RB = D / 16;                  // 16-row blocks of U
    constexpr int WAVES = TB / 64;
    const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1], b2 = 1u << op.p[2], b3 = 1u << op.p[3];
    const uint32_t b4 = K > 4 ? 1u << op.p[4] : 0u;
    const uint32_t tm = b0 | b1 | b2 | b3 | b4;
    const uint32_t groups = T >> K;
    const uint32_t cb2 = groups / 32;          // pairs of 16-column blocks
    const uint32_t tasks = RB * cb2;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t lr = lane & 15u, lk = lane >> 4;
    acc_t yr[MAXT][2], yi[MAXT][2];
    uint32_t taskId[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        taskId[t] = wave + (uint32_t)t * WAVES;
#pragma unroll
        for (int c = 0; c < 2; c++) yr[t][c] = yi[t][c] = acc_t{0, 0, 0, 0};
        if (taskId[t] >= tasks) continue;
        const uint32_t rb = taskId[t] % RB, cbase = (taskId[t] / RB) * 2;
        const uint32_t urow = rb * 16 + lr;               // row of U this lane feeds
        uint32_t gb[2];
#pragma unroll
        for (int c = 0; c < 2; c++) gb[c] = ins0_mask32((cbase + (uint32_t)c) * 16 + lr, tm);
        const cx<double>* mrow = M + urow * (uint32_t)D;
        const uint32_t sw = urow & (uint32_t)(D - 1);
        // not unrolled: the MFMAs of the other waves on the SIMD cover the
        // operand latency, and the loops stay clear of register spills
#pragma unroll 1
        for (uint32_t ks = 0; ks < (uint32_t)(D / 4); ks++) {
            const uint32_t col = ks * 4 + lk;             // column of U = row of X
            const cx<double> u = mrow[col ^ sw];
            const uint32_t off = mat_offset<K>(col, b0, b1, b2, b3, b4);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const cx<double> x = s[gb[c] | off];
                yr[t][c] = MfmaAcc<double>::step(u.x, x.x, yr[t][c]);
                yi[t][c] = MfmaAcc<double>::step(u.y, x.y, yi[t][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const acc_t p1 = yr[t][c], p2 = yi[t][c];
            yr[t][c] = p1 - p2;
            yi[t][c] = -(p1 + p2);
        }
#pragma unroll 1
        for (uint32_t ks = 0; ks < (uint32_t)(D / 4); ks++) {
            const uint32_t col = ks * 4 + lk;
            const cx<double> u = mrow[col ^ sw];
            const double us = u.x + u.y;
            const uint32_t off = mat_offset<K>(col, b0, b1, b2, b3, b4);
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const cx<double> x = s[gb[c] | off];
                yi[t][c] = MfmaAcc<double>::step(us, x.x + x.y, yi[t][c]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        if (taskId[t] >= tasks) continue;
        const uint32_t rb = taskId[t] % RB, cbase = (taskId[t] / RB) * 2;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t gbc = ins0_mask32((cbase + (uint32_t)c) * 16 + lr, tm);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // accumulator element j of this lane: row 4 j + lane / 16 of U
                const uint32_t row = rb * 16 + 4u * (uint32_t)j + lk;
                s[gbc | mat_offset<K>(row, b0, b1, b2, b3, b4)] = mk<double>(yr[t][c][j], yi[t][c][j]);
            }
        }
    }
}

// Kernel variants: VAR 0 carries the simple LDS ops (and, with SW, register
// sweeps); VAR 1 adds VALU dense gates and two-qubit depolarising; VAR 2 has
// MFMA dense gates (matrices in LDS for the whole pass) instead of VALU ones,
// VAR 3 the same with each matrix staged in LDS just before its item (passes
// whose matrices exceed the LDS area).  Each instantiation stays within 128 VGPRs so
// two 64 KiB workgroups share a CU.
// lmat: the pass's LDS matrix area (VAR 2: every MFMA item has an offset in it)
// Simple items over n pairs (i0, i1) = idx(j) or n single amplitudes l = j,
// j = tid, tid + TB, ...  With KB pairs (KB amplitudes) per thread known at
// compile time and n = KB * TB, the LDS loads of up to GM pairs (2 GM
// amplitudes) are issued before their updates and stores, so the item pays the
// LDS latency once per batch instead of once per pair; the slots of different
// j are distinct, so the early loads read nothing the stores change.  GM comes
// from the kernel, as many as its registers allow: 4 in the plain FP64 kernel,
// 2 in the dense FP64 one, 1 in FP32 (which keeps the 32 KiB-tile kernel
// within 64 VGPRs, four workgroups per CU).  Otherwise (KB < 2, partial tiles)
// one pair at a time.
template <typename R, int TB, int KB, int GM, typename C, typename Idx, typename Upd>
__device__ __forceinline__ void tile_pairs(cx<R>* s, uint32_t n, Idx idx, Upd upd) {
    const uint32_t tid = threadIdx.x;
    if constexpr (KB >= 2 && GM >= 1) {
        constexpr int G = KB < GM ? KB : GM;
        if (n == (uint32_t)KB * TB) {
#pragma unroll
            for (int k0 = 0; k0 < KB; k0 += G) {
                uint32_t i0[G], i1[G];
                cx<C> v0[G], v1[G];
#pragma unroll
                for (int k = 0; k < G; k++) {
                    idx(tid + (uint32_t)(k0 + k) * TB, i0[k], i1[k]);
                    v0[k] = up<C>(s[i0[k]]);
                    v1[k] = up<C>(s[i1[k]]);
                }
#pragma unroll
                for (int k = 0; k < G; k++) {
                    upd(i0[k], i1[k], v0[k], v1[k]);
                    s[i0[k]] = dn<R>(v0[k]);
                    s[i1[k]] = dn<R>(v1[k]);
                }
            }
            return;
        }
    }
    for (uint32_t j = tid; j < n; j += TB) {
        uint32_t i0, i1;
        idx(j, i0, i1);
        cx<C> v0 = up<C>(s[i0]), v1 = up<C>(s[i1]);
        upd(i0, i1, v0, v1);
        s[i0] = dn<R>(v0);
        s[i1] = dn<R>(v1);
    }
}

template <typename R, int TB, int KS, int GM, typename C, typename Upd>
__device__ __forceinline__ void tile_singles(cx<R>* s, uint32_t T, Upd upd) {
    const uint32_t tid = threadIdx.x;
    if constexpr (KS >= 2 && GM >= 1) {
        constexpr int G = KS < GM ? KS : GM;
        if (T == (uint32_t)KS * TB) {
#pragma unroll
            for (int k0 = 0; k0 < KS; k0 += G) {
                cx<C> v[G];
#pragma unroll
                for (int k = 0; k < G; k++) v[k] = up<C>(s[tid + (uint32_t)(k0 + k) * TB]);
#pragma unroll
                for (int k = 0; k < G; k++) {
                    upd(tid + (uint32_t)(k0 + k) * TB, v[k]);
                    s[tid + (uint32_t)(k0 + k) * TB] = dn<R>(v[k]);
                }
            }
            return;
        }
    }
    for (uint32_t l = tid; l < T; l += TB) {
        cx<C> v = up<C>(s[l]);
        upd(l, v);
        s[l] = dn<R>(v);
    }
}

// Quartets (i, i | b0, i | b1, i | b0 | b1) of two channel bits, GM at a time
// as tile_pairs does for pairs (n = KB * TB quartets).
template <typename R, int TB, int KB, int GM, typename C, typename Upd>
__device__ __forceinline__ void tile_quads(cx<R>* s, uint32_t n, uint32_t b0, uint32_t b1, Upd upd) {
    const uint32_t tid = threadIdx.x;
    static_assert(KB >= 2 && GM >= 1, "batched quartets only");
    constexpr int G = KB < GM ? KB : GM;
#pragma unroll
    for (int k0 = 0; k0 < KB; k0 += G) {
        uint32_t i[G];
        cx<C> v[G][4];
#pragma unroll
        for (int k = 0; k < G; k++) {
            i[k] = ins0_mask32(tid + (uint32_t)(k0 + k) * TB, b0 | b1);
            v[k][0] = up<C>(s[i[k]]);
            v[k][1] = up<C>(s[i[k] | b0]);
            v[k][2] = up<C>(s[i[k] | b1]);
            v[k][3] = up<C>(s[i[k] | b0 | b1]);
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
            upd(v[k]);
            s[i[k]] = dn<R>(v[k][0]);
            s[i[k] | b0] = dn<R>(v[k][1]);
            s[i[k] | b1] = dn<R>(v[k][2]);
            s[i[k] | b0 | b1] = dn<R>(v[k][3]);
        }
    }
    (void)n;
}

// ROWS: amplitudes per thread of a full tile for the batched simple items; GP:
// pairs per batch (singles 2 GP, density-matrix quartets GP / 2 in FP64; 0:
// one at a time)
template <typename R, int TB, int VAR, int MAXT = 1, typename C = R, int ROWS = 0, int GP = 0>
__device__ __forceinline__ void tile_apply(cx<R>* s, uint32_t T, uint64_t base, const TileOp& op,
                           const cx<C>* __restrict__ mats, const R* lmat = nullptr) {
    const uint32_t tid = threadIdx.x;
    // the density-matrix channels batch in FP64 only (FP32: registers)
    constexpr int GPC = sizeof(C) == 8 ? GP : 0;
    switch (op.kind) {
    case OP_MAT1: {
        if ((base & op.gmask) != op.gmask) return;           // an outside control is 0
        const uint32_t tb = 1u << op.p[0];
        const uint32_t lc = op.lmask;
        const uint32_t nIts = T >> (1 + __builtin_popcount(lc));
        const cx<C> g00 = ldc<C>(op.f, 0), g01 = ldc<C>(op.f, 1), g10 = ldc<C>(op.f, 2), g11 = ldc<C>(op.f, 3);
        if (lc == 0) {
            const unsigned t = op.p[0];
            tile_pairs<R, TB, ROWS / 2, GP, C>(
                s, nIts, [&](uint32_t j, uint32_t& i0, uint32_t& i1) { i0 = ins0_32(j, t); i1 = i0 | tb; },
                [&](uint32_t, uint32_t, cx<C>& v0, cx<C>& v1) {
                    const cx<C> a = v0;
                    v0 = cfma(cmul(g00, a), g01, v1);
                    v1 = cfma(cmul(g10, a), g11, v1);
                });
            return;
        }
        for (uint32_t j = tid; j < nIts; j += TB) {
            uint32_t i0 = ins0_mask32(j, lc | tb) | lc;
            cx<C> v0 = up<C>(s[i0]), v1 = up<C>(s[i0 | tb]);
            s[i0] = dn<R>(cfma(cmul(g00, v0), g01, v1));
            s[i0 | tb] = dn<R>(cfma(cmul(g10, v0), g11, v1));
        }
        return;
    }
    case OP_SWAP: {
        const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1];
        for (uint32_t j = tid; j < (T >> 2); j += TB) {
            uint32_t i = ins0_mask32(j, b0 | b1);
            cx<R> x = s[i | b0], y = s[i | b1];
            s[i | b0] = y;
            s[i | b1] = x;
        }
        return;
    }
    case OP_PAULI: {
        const cx<C> th = ldc<C>(op.f, 0), fe = ldc<C>(op.f, 1);
        const cx<C> fo = mk<C>(-fe.x, -fe.y);
        const uint32_t bp = par64(base & op.gmask);
        const uint32_t lxy = op.lmask, lyz = op.lmask2;
        if (lxy == 0) {
            tile_singles<R, TB, ROWS, 2 * GP, C>(s, T, [&](uint32_t l, cx<C>& v) {
                const cx<C> f = (bp ^ par32(l & lyz)) ? fo : fe;
                v = cfma(cmul(th, v), f, v);
            });
            return;
        }
        const unsigned hb = 31u - (unsigned)__builtin_clz(lxy);
        tile_pairs<R, TB, ROWS / 2, GP, C>(
            s, T >> 1, [&](uint32_t j, uint32_t& i0, uint32_t& i1) { i0 = ins0_32(j, hb); i1 = i0 ^ lxy; },
            [&](uint32_t i0, uint32_t i1, cx<C>& v0, cx<C>& v1) {
                const cx<C> f0 = (bp ^ par32(i0 & lyz)) ? fo : fe;
                const cx<C> f1 = (bp ^ par32(i1 & lyz)) ? fo : fe;
                const cx<C> a = v0;
                v0 = cfma(cmul(th, a), f1, v1);
                v1 = cfma(cmul(th, v1), f0, a);
            });
        return;
    }
    case OP_PHASE: {
        const cx<C> f0 = ldc<C>(op.f, 0), f1 = ldc<C>(op.f, 1);
        const uint32_t bp = par64(base & op.gmask);
        const uint32_t lm = op.lmask;
        tile_singles<R, TB, ROWS, 2 * GP, C>(s, T, [&](uint32_t l, cx<C>& v) { v = cmul(v, (bp ^ par32(l & lm)) ? f1 : f0); });
        return;
    }
    case OP_MATK: {
        if constexpr (VAR == 0) return;
        const cx<C>* mat = mats + op.mat_off;
        if (op.lmask2 == 1u) {        // host chose the MFMA form (full tile, k = 3, 4 or 5)
            if constexpr (VAR >= 2) {
                static_assert(sizeof(C) == sizeof(R), "MFMA tiles compute in their storage precision");
                // the matrix is in LDS: staged for the whole pass (VAR 2) or for this item (VAR 3)
                const R* M = lmat + op.lmask;
                if constexpr (tile_mfma3_form<R, TB>(4)) {
                    if (op.k >= 4) {          // three-product form, complex U in LDS
                        const cx<double>* Mc = reinterpret_cast<const cx<double>*>(M);
                        if (op.k == 4) tile_matk_mfma3<4, TB, MAXT>(s, T, op, Mc);
                        else tile_matk_mfma3<5, TB, MAXT>(s, T, op, Mc);
                        return;
                    }
                }
                if (op.k == 3) tile_matk_mfma<R, 3, TB, MAXT>(s, T, op, M);
                else if (op.k == 4) tile_matk_mfma<R, 4, TB, MAXT>(s, T, op, M);
                else tile_matk_mfma<R, 5, TB, MAXT>(s, T, op, M);
            }
            return;
        }
        if constexpr (VAR >= 2) {
            // the MFMA variants also carry two-target gates (VALU, in place)
            if (op.k == 2) tile_matk<R, 2, TB, C>(s, T, op, mat);
            return;
        }
        switch (op.k) {
        case 1: tile_matk<R, 1, TB, C>(s, T, op, mat); break;
        case 2: tile_matk<R, 2, TB, C>(s, T, op, mat); break;
        case 3: tile_matk<R, 3, TB, C>(s, T, op, mat); break;
        case 4: tile_matk<R, 4, TB, C>(s, T, op, mat); break;
        case 5: tile_matk<R, 5, TB, C>(s, T, op, mat); break;
        default: break;
        }
        return;
    }
    case OP_DEPHASE1: {
        const C fac = (C)op.f[0];
        const uint32_t cq = op.p[0], ca = op.p[1];
        if constexpr (GPC > 0) {
            tile_singles<R, TB, ROWS, 2 * GPC, C>(s, T, [&](uint32_t l, cx<C>& v) {
                if (chan_bit(l, base, cq) ^ chan_bit(l, base, ca)) v = cscale(v, fac);
            });
            return;
        }
        for (uint32_t l = tid; l < T; l += TB)
            if (chan_bit(l, base, cq) ^ chan_bit(l, base, ca)) s[l] = dn<R>(cscale(up<C>(s[l]), fac));
        return;
    }
    case OP_DEPHASE2: {
        const C fac = (C)op.f[0];
        if constexpr (GPC > 0) {
            const uint32_t q0 = op.p[0], q1 = op.p[1], q2 = op.p[2], q3 = op.p[3];
            tile_singles<R, TB, ROWS, 2 * GPC, C>(s, T, [&](uint32_t l, cx<C>& v) {
                if ((chan_bit(l, base, q0) ^ chan_bit(l, base, q1)) | (chan_bit(l, base, q2) ^ chan_bit(l, base, q3)))
                    v = cscale(v, fac);
            });
            return;
        }
        for (uint32_t l = tid; l < T; l += TB) {
            uint32_t f = (chan_bit(l, base, op.p[0]) ^ chan_bit(l, base, op.p[1])) |
                         (chan_bit(l, base, op.p[2]) ^ chan_bit(l, base, op.p[3]));
            if (f) s[l] = dn<R>(cscale(up<C>(s[l]), fac));
        }
        return;
    }
    case OP_DEPOL1: {
        const C c1 = (C)op.f[0], c2 = (C)op.f[1], c3 = (C)op.f[2];
        const uint32_t bq = 1u << op.p[0], ba = 1u << op.p[1];
        if constexpr (GPC / 2 > 0 && ROWS / 4 >= 2) {
            if (T == (uint32_t)ROWS * TB) {
                tile_quads<R, TB, ROWS / 4, GPC / 2, C>(s, T >> 2, bq, ba, [&](cx<C> (&v)[4]) {
                    const cx<C> v00 = v[0], v11 = v[3];
                    v[0] = mk<C>(c2 * v00.x + c1 * v11.x, c2 * v00.y + c1 * v11.y);
                    v[1] = cscale(v[1], c3);
                    v[2] = cscale(v[2], c3);
                    v[3] = mk<C>(c1 * v00.x + c2 * v11.x, c1 * v00.y + c2 * v11.y);
                });
                return;
            }
        }
        for (uint32_t j = tid; j < (T >> 2); j += TB) {
            uint32_t i = ins0_mask32(j, bq | ba);
            cx<C> v00 = up<C>(s[i]), v01 = up<C>(s[i | bq]), v10 = up<C>(s[i | ba]), v11 = up<C>(s[i | bq | ba]);
            s[i] = dn<R>(mk<C>(c2 * v00.x + c1 * v11.x, c2 * v00.y + c1 * v11.y));
            s[i | bq] = dn<R>(cscale(v01, c3));
            s[i | ba] = dn<R>(cscale(v10, c3));
            s[i | bq | ba] = dn<R>(mk<C>(c1 * v00.x + c2 * v11.x, c1 * v00.y + c2 * v11.y));
        }
        return;
    }
    case OP_DAMP: {
        const C prob = (C)op.f[0], c1 = (C)op.f[1], c2 = (C)op.f[2];
        const uint32_t bq = 1u << op.p[0], ba = 1u << op.p[1];
        if constexpr (GPC / 2 > 0 && ROWS / 4 >= 2) {
            if (T == (uint32_t)ROWS * TB) {
                tile_quads<R, TB, ROWS / 4, GPC / 2, C>(s, T >> 2, bq, ba, [&](cx<C> (&v)[4]) {
                    const cx<C> v00 = v[0], v11 = v[3];
                    v[0] = mk<C>(v00.x + prob * v11.x, v00.y + prob * v11.y);
                    v[1] = cscale(v[1], c1);
                    v[2] = cscale(v[2], c1);
                    v[3] = cscale(v11, c2);
                });
                return;
            }
        }
        for (uint32_t j = tid; j < (T >> 2); j += TB) {
            uint32_t i = ins0_mask32(j, bq | ba);
            cx<C> v00 = up<C>(s[i]), v01 = up<C>(s[i | bq]), v10 = up<C>(s[i | ba]), v11 = up<C>(s[i | bq | ba]);
            s[i] = dn<R>(mk<C>(v00.x + prob * v11.x, v00.y + prob * v11.y));
            s[i | bq] = dn<R>(cscale(v01, c1));
            s[i | ba] = dn<R>(cscale(v10, c1));
            s[i | bq | ba] = dn<R>(cscale(v11, c2));
        }
        return;
    }
    case OP_DEPOL2: {
        if constexpr (VAR == 0) return;
        const C c1 = (C)op.f[0], c2 = (C)op.f[1], sc = (C)op.f[2];
        const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1], b2 = 1u << op.p[2], b3 = 1u << op.p[3];
        const uint32_t mk4 = b0 | b1 | b2 | b3;
        for (uint32_t j = tid; j < (T >> 4); j += TB) {
            uint32_t i = ins0_mask32(j, mk4);
            // block-diagonal quartet: row bits equal column bits on both qubits
            const uint32_t d0 = i, d1 = i | b0 | b2, d2 = i | b1 | b3, d3 = i | mk4;
            cx<C> v0 = up<C>(s[d0]), v1 = up<C>(s[d1]), v2 = up<C>(s[d2]), v3 = up<C>(s[d3]);
            cx<C> sum = cadd(cadd(v0, v1), cadd(v2, v3));
            s[d0] = dn<R>(mk<C>(c1 * v0.x + c2 * sum.x, c1 * v0.y + c2 * sum.y));
            s[d1] = dn<R>(mk<C>(c1 * v1.x + c2 * sum.x, c1 * v1.y + c2 * sum.y));
            s[d2] = dn<R>(mk<C>(c1 * v2.x + c2 * sum.x, c1 * v2.y + c2 * sum.y));
            s[d3] = dn<R>(mk<C>(c1 * v3.x + c2 * sum.x, c1 * v3.y + c2 * sum.y));
#pragma unroll
            for (int l = 0; l < 16; l++) {
                if (l == 0 || l == 5 || l == 10 || l == 15) continue;
                uint32_t x = i | ((l & 1) ? b0 : 0u) | ((l & 2) ? b1 : 0u) | ((l & 4) ? b2 : 0u) | ((l & 8) ? b3 : 0u);
                s[x] = dn<R>(cscale(up<C>(s[x]), sc));
            }
        }
        return;
    }
    default:
        return;
    }
}

}  // namespace tile
}  // namespace fsa
