// kernels_tile.hip -- fused multi-operation pass over LDS tiles.
//
// A pass is defined by a set of m "tile qubits" (sorted global positions, the
// low ones first so each run of consecutive tile indices is contiguous in
// HBM).  Workgroup b owns the 2^m amplitudes whose other bits spell b: it
// gathers them into LDS (one coalesced read), applies every operation of the
// pass whose non-diagonal qubits are tile qubits, and scatters them back (one
// coalesced write).  Controls and Z-type factors on qubits outside the tile
// are constant over a workgroup and are folded into a per-tile test or sign.
// The arithmetic per amplitude is the same as the streaming kernels'.
#include "launch.hpp"

#include <algorithm>

namespace fsa {

namespace {

// bit value of a channel qubit: code bit 8 set -> local position in the tile,
// otherwise a global qubit index outside the tile (fixed for the workgroup)
__device__ __forceinline__ uint32_t chan_bit(uint32_t l, uint64_t base, uint32_t code) {
    return (code & 0x100u) ? (l >> (code & 0xffu)) & 1u : (uint32_t)((base >> code) & 1ULL);
}

template <typename R>
__device__ __forceinline__ cx<R> ldc(const double* f, int i) { return mk<R>((R)f[2 * i], (R)f[2 * i + 1]); }

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

template <typename R, int K, int TB>
__device__ __forceinline__ void tile_matk(cx<R>* s, uint32_t T, const TileOp& op, const cx<R>* __restrict__ mat) {
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
    constexpr int OUTMAX = sizeof(R) == 8 ? 8 : 16;
    constexpr int OUT = D < OUTMAX ? D : OUTMAX;
    // whole groups per thread, written in place; a full tile has >= TB groups only when D <= OUT
    if constexpr (D <= OUT) if (groups >= (uint32_t)TB) {
        for (uint32_t g = tid; g < groups; g += TB) {
            const cx<R>* m = mat;
            const uint32_t b = ins0_mask32(g, tm);
            cx<R> acc[D];
#pragma unroll
            for (int o = 0; o < D; o++) acc[o] = mk<R>(0, 0);
            // input loop kept rolled: D accumulators and one matrix column live at a time
#pragma unroll 1
            for (int l = 0; l < D; l++) {
                cx<R> x = s[b | FSA_OFF((uint32_t)l)];
#pragma unroll
                for (int o = 0; o < D; o++) acc[o] = cfma(acc[o], m[o * D + l], x);
            }
#pragma unroll
            for (int o = 0; o < D; o++) s[b | FSA_OFF((uint32_t)o)] = acc[o];
        }
        return;
    }
    // split mode (one group per thread, tpg threads per group, outs consecutive
    // outputs each): all inputs are read before the barrier, outputs written after
    const uint32_t byThreads = TB / groups;
    const uint32_t tpg = byThreads < (uint32_t)(D / OUT) ? byThreads : (uint32_t)(D / OUT);
    const uint32_t outs = D / tpg;
    const uint32_t g = tid & (groups - 1), sub = tid / groups;
    const bool active = sub < tpg;
    const uint32_t r0 = sub * outs;
    const uint32_t b = ins0_mask32(g, tm);
    cx<R> acc[OUT];
#pragma unroll
    for (int o = 0; o < OUT; o++) acc[o] = mk<R>(0, 0);
    if (groups >= 64 && outs == (uint32_t)OUT) {
        // a wave spans one `sub`, so its output rows are wave-uniform: read the
        // matrix through the scalar unit instead of one vector load per lane
        const uint32_t r0u = __builtin_amdgcn_readfirstlane(r0);
        const cx<R>* mrow = mat + (size_t)r0u * D;
        if (active) {
#pragma unroll 1
            for (int l = 0; l < D; l++) {
                cx<R> x = s[b | FSA_OFF((uint32_t)l)];
#pragma unroll
                for (int o = 0; o < OUT; o++) acc[o] = cfma(acc[o], mrow[o * D + l], x);
            }
        }
    } else if (active) {
#pragma unroll 1
        for (int l = 0; l < D; l++) {
            cx<R> x = s[b | FSA_OFF((uint32_t)l)];
#pragma unroll
            for (int o = 0; o < OUT; o++)
                if ((uint32_t)o < outs) acc[o] = cfma(acc[o], mat[(r0 + o) * D + l], x);
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int o = 0; o < OUT; o++)
            if ((uint32_t)o < outs) s[b | FSA_OFF(r0 + o)] = acc[o];
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

// tasks per wave for a full tile: FP64 (4096 amplitudes) 1, FP32 (8192) 2
template <typename R, int K, int TB>
__device__ __forceinline__ void tile_matk_mfma(cx<R>* s, uint32_t T, const TileOp& op, const R* __restrict__ M) {
    typedef typename MfmaAcc<R>::T acc_t;
    constexpr int D = 1 << K, D2 = 2 * D;
    constexpr int RB = D2 / 16;                 // 16-row blocks of the real matrix
    constexpr int WAVES = TB / 64;
    constexpr int MAXT = sizeof(R) == 8 ? 1 : 2;
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
        const R* arow = M + (size_t)(rb * 16 + lr) * D2;
        // B-fragment group bases of the four column blocks for this lane
        uint32_t gb[4];
#pragma unroll
        for (int c = 0; c < 4; c++) gb[c] = ins0_mask32((cbase + (uint32_t)c) * 16 + lr, tm);
        const R* sr = reinterpret_cast<const R*>(s);
        // four k-steps per iteration: their A loads and LDS reads issue together
#pragma unroll 4
        for (int ks = 0; ks < D2 / 4; ks++) {
            const R a = arow[ks * 4 + lk];
            const uint32_t r = (uint32_t)ks * 4 + lk;          // real row of X this lane feeds
            const uint32_t off = mat_offset<K>(r & (D - 1), b0, b1, b2, b3, b4);
            const uint32_t half = r < (uint32_t)D ? 0u : 1u;  // real or imaginary part only
#pragma unroll
            for (int c = 0; c < 4; c++) acc[t][c] = MfmaAcc<R>::step(a, sr[2 * (gb[c] | off) + half], acc[t][c]);
        }
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

// Kernel variants: VAR 0 carries the simple LDS ops (and, with SW, register
// sweeps); VAR 1 adds VALU dense gates and two-qubit depolarising; VAR 2 has
// MFMA dense gates instead of VALU ones.  Each instantiation stays within 128
// VGPRs so two 64 KiB workgroups share a CU.
template <typename R, int TB, int VAR>
__device__ __forceinline__ void tile_apply(cx<R>* s, uint32_t T, uint64_t base, const TileOp& op,
                           const cx<R>* __restrict__ mats) {
    const uint32_t tid = threadIdx.x;
    switch (op.kind) {
    case OP_MAT1: {
        if ((base & op.gmask) != op.gmask) return;           // an outside control is 0
        const uint32_t tb = 1u << op.p[0];
        const uint32_t lc = op.lmask;
        const uint32_t nIts = T >> (1 + __builtin_popcount(lc));
        const cx<R> g00 = ldc<R>(op.f, 0), g01 = ldc<R>(op.f, 1), g10 = ldc<R>(op.f, 2), g11 = ldc<R>(op.f, 3);
        for (uint32_t j = tid; j < nIts; j += TB) {
            uint32_t i0 = ins0_mask32(j, lc | tb) | lc;
            cx<R> v0 = s[i0], v1 = s[i0 | tb];
            s[i0] = cfma(cmul(g00, v0), g01, v1);
            s[i0 | tb] = cfma(cmul(g10, v0), g11, v1);
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
        const cx<R> th = ldc<R>(op.f, 0), fe = ldc<R>(op.f, 1);
        const cx<R> fo = mk<R>(-fe.x, -fe.y);
        const uint32_t bp = par64(base & op.gmask);
        const uint32_t lxy = op.lmask, lyz = op.lmask2;
        if (lxy == 0) {
            for (uint32_t l = tid; l < T; l += TB) {
                cx<R> v = s[l];
                cx<R> f = (bp ^ par32(l & lyz)) ? fo : fe;
                s[l] = cfma(cmul(th, v), f, v);
            }
            return;
        }
        const unsigned hb = 31u - (unsigned)__builtin_clz(lxy);
        for (uint32_t j = tid; j < (T >> 1); j += TB) {
            uint32_t i0 = ins0_32(j, hb), i1 = i0 ^ lxy;
            cx<R> v0 = s[i0], v1 = s[i1];
            cx<R> f0 = (bp ^ par32(i0 & lyz)) ? fo : fe;
            cx<R> f1 = (bp ^ par32(i1 & lyz)) ? fo : fe;
            s[i0] = cfma(cmul(th, v0), f1, v1);
            s[i1] = cfma(cmul(th, v1), f0, v0);
        }
        return;
    }
    case OP_PHASE: {
        const cx<R> f0 = ldc<R>(op.f, 0), f1 = ldc<R>(op.f, 1);
        const uint32_t bp = par64(base & op.gmask);
        for (uint32_t l = tid; l < T; l += TB)
            s[l] = cmul(s[l], (bp ^ par32(l & op.lmask)) ? f1 : f0);
        return;
    }
    case OP_MATK: {
        if constexpr (VAR == 0) return;
        const cx<R>* mat = mats + op.mat_off;
        if (op.lmask2 == 1u) {        // host chose the MFMA form (full tile, k = 3, 4 or 5)
            if constexpr (VAR == 2) {
                const R* M = reinterpret_cast<const R*>(mat);
                if (op.k == 3) tile_matk_mfma<R, 3, TB>(s, T, op, M);
                else if (op.k == 4) tile_matk_mfma<R, 4, TB>(s, T, op, M);
                else tile_matk_mfma<R, 5, TB>(s, T, op, M);
            }
            return;
        }
        if constexpr (VAR == 2) return;
        switch (op.k) {
        case 1: tile_matk<R, 1, TB>(s, T, op, mat); break;
        case 2: tile_matk<R, 2, TB>(s, T, op, mat); break;
        case 3: tile_matk<R, 3, TB>(s, T, op, mat); break;
        case 4: tile_matk<R, 4, TB>(s, T, op, mat); break;
        case 5: tile_matk<R, 5, TB>(s, T, op, mat); break;
        default: break;
        }
        return;
    }
    case OP_DEPHASE1: {
        const R fac = (R)op.f[0];
        const uint32_t cq = op.p[0], ca = op.p[1];
        for (uint32_t l = tid; l < T; l += TB)
            if (chan_bit(l, base, cq) ^ chan_bit(l, base, ca)) s[l] = cscale(s[l], fac);
        return;
    }
    case OP_DEPHASE2: {
        const R fac = (R)op.f[0];
        for (uint32_t l = tid; l < T; l += TB) {
            uint32_t f = (chan_bit(l, base, op.p[0]) ^ chan_bit(l, base, op.p[1])) |
                         (chan_bit(l, base, op.p[2]) ^ chan_bit(l, base, op.p[3]));
            if (f) s[l] = cscale(s[l], fac);
        }
        return;
    }
    case OP_DEPOL1: {
        const R c1 = (R)op.f[0], c2 = (R)op.f[1], c3 = (R)op.f[2];
        const uint32_t bq = 1u << op.p[0], ba = 1u << op.p[1];
        for (uint32_t j = tid; j < (T >> 2); j += TB) {
            uint32_t i = ins0_mask32(j, bq | ba);
            cx<R> v00 = s[i], v01 = s[i | bq], v10 = s[i | ba], v11 = s[i | bq | ba];
            s[i] = mk<R>(c2 * v00.x + c1 * v11.x, c2 * v00.y + c1 * v11.y);
            s[i | bq] = cscale(v01, c3);
            s[i | ba] = cscale(v10, c3);
            s[i | bq | ba] = mk<R>(c1 * v00.x + c2 * v11.x, c1 * v00.y + c2 * v11.y);
        }
        return;
    }
    case OP_DAMP: {
        const R prob = (R)op.f[0], c1 = (R)op.f[1], c2 = (R)op.f[2];
        const uint32_t bq = 1u << op.p[0], ba = 1u << op.p[1];
        for (uint32_t j = tid; j < (T >> 2); j += TB) {
            uint32_t i = ins0_mask32(j, bq | ba);
            cx<R> v00 = s[i], v01 = s[i | bq], v10 = s[i | ba], v11 = s[i | bq | ba];
            s[i] = mk<R>(v00.x + prob * v11.x, v00.y + prob * v11.y);
            s[i | bq] = cscale(v01, c1);
            s[i | ba] = cscale(v10, c1);
            s[i | bq | ba] = cscale(v11, c2);
        }
        return;
    }
    case OP_DEPOL2: {
        if constexpr (VAR == 0) return;
        const R c1 = (R)op.f[0], c2 = (R)op.f[1], sc = (R)op.f[2];
        const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1], b2 = 1u << op.p[2], b3 = 1u << op.p[3];
        const uint32_t mk4 = b0 | b1 | b2 | b3;
        for (uint32_t j = tid; j < (T >> 4); j += TB) {
            uint32_t i = ins0_mask32(j, mk4);
            // block-diagonal quartet: row bits equal column bits on both qubits
            const uint32_t d0 = i, d1 = i | b0 | b2, d2 = i | b1 | b3, d3 = i | mk4;
            cx<R> v0 = s[d0], v1 = s[d1], v2 = s[d2], v3 = s[d3];
            cx<R> sum = cadd(cadd(v0, v1), cadd(v2, v3));
            s[d0] = mk<R>(c1 * v0.x + c2 * sum.x, c1 * v0.y + c2 * sum.y);
            s[d1] = mk<R>(c1 * v1.x + c2 * sum.x, c1 * v1.y + c2 * sum.y);
            s[d2] = mk<R>(c1 * v2.x + c2 * sum.x, c1 * v2.y + c2 * sum.y);
            s[d3] = mk<R>(c1 * v3.x + c2 * sum.x, c1 * v3.y + c2 * sum.y);
#pragma unroll
            for (int l = 0; l < 16; l++) {
                if (l == 0 || l == 5 || l == 10 || l == 15) continue;
                uint32_t x = i | ((l & 1) ? b0 : 0u) | ((l & 2) ? b1 : 0u) | ((l & 4) ? b2 : 0u) | ((l & 8) ? b3 : 0u);
                s[x] = cscale(s[x], sc);
            }
        }
        return;
    }
    default:
        return;
    }
}

// ---- register sweeps ------------------------------------------------------
// A sweep gives each thread G = 2^W amplitudes of the tile: tile-local indices
// gb | dep(i), where gb has zeros at the W sweep bits and dep(i) places bit j
// of i at sweep bit j.  Every micro-op of the sweep acts only on sweep bits
// non-diagonally, so it runs on registers; the sweep costs one LDS round trip
// (none when it is fused with the HBM load or store).  Sweep-local bit
// positions are template arguments so register indices stay compile-time.

__device__ __forceinline__ uint32_t sw_bit(uint32_t code, int i, uint32_t gb, uint64_t base) {
    // code: 0x200|j sweep bit j; 0x100|p tile bit p outside the sweep; else a global qubit outside the tile
    return (code & 0x200u) ? ((uint32_t)i >> (code & 0xffu)) & 1u
         : (code & 0x100u) ? (gb >> (code & 0xffu)) & 1u
                           : (uint32_t)((base >> code) & 1ULL);
}

template <typename R, int W, int J>
__device__ __forceinline__ void sw_mat1(cx<R>* v, const TileOp& o) {
    const cx<R> g00 = ldc<R>(o.f, 0), g01 = ldc<R>(o.f, 1), g10 = ldc<R>(o.f, 2), g11 = ldc<R>(o.f, 3);
    const uint32_t cs = o.lmask;
#pragma unroll
    for (int i = 0; i < (1 << W); i++) {
        if (i & (1 << J)) continue;
        if (((uint32_t)i & cs) != cs) continue;
        cx<R> a0 = v[i], a1 = v[i | (1 << J)];
        v[i] = cfma(cmul(g00, a0), g01, a1);
        v[i | (1 << J)] = cfma(cmul(g10, a0), g11, a1);
    }
}

template <typename R, int W, int MXY>
__device__ __forceinline__ void sw_pauli(cx<R>* v, const TileOp& o, uint32_t gpar) {
    const cx<R> th = ldc<R>(o.f, 0), fe = ldc<R>(o.f, 1);
    const cx<R> fo = mk<R>(-fe.x, -fe.y);
    const uint32_t syz = o.lmask2;
#pragma unroll
    for (int i = 0; i < (1 << W); i++) {
        if (MXY == 0) {
            cx<R> a = v[i];
            cx<R> f = (gpar ^ par32((uint32_t)i & syz)) ? fo : fe;
            v[i] = cfma(cmul(th, a), f, a);
            continue;
        }
        const int j = i ^ MXY;
        if (j < i) continue;
        cx<R> a0 = v[i], a1 = v[j];
        cx<R> f0 = (gpar ^ par32((uint32_t)i & syz)) ? fo : fe;
        cx<R> f1 = (gpar ^ par32((uint32_t)j & syz)) ? fo : fe;
        v[i] = cfma(cmul(th, a0), f1, a1);
        v[j] = cfma(cmul(th, a1), f0, a0);
    }
}

template <typename R, int W, int A, int B>
__device__ __forceinline__ void sw_swap(cx<R>* v) {
#pragma unroll
    for (int i = 0; i < (1 << W); i++) {
        if (!(i & (1 << A)) || (i & (1 << B))) continue;
        const int j = i ^ (1 << A) ^ (1 << B);
        cx<R> t = v[i];
        v[i] = v[j];
        v[j] = t;
    }
}

// dense gate on sweep bits B0 < B1 (< B2 < B3); the host orders the matrix so
// matrix bit j acts on the j-th smallest sweep bit
template <typename R, int W, int K, int B0, int B1, int B2, int B3>
__device__ __forceinline__ void sw_matk(cx<R>* v, const cx<R>* __restrict__ m) {
    constexpr int D = 1 << K;
    constexpr int MSK = (1 << B0) | (K > 1 ? 1 << B1 : 0) | (K > 2 ? 1 << B2 : 0) | (K > 3 ? 1 << B3 : 0);
#pragma unroll
    for (int b = 0; b < (1 << W); b++) {
        if (b & MSK) continue;
        cx<R> x[D];
#pragma unroll
        for (int l = 0; l < D; l++) {
            int idx = b | ((l & 1) << B0);
            if (K > 1) idx |= ((l >> 1) & 1) << B1;
            if (K > 2) idx |= ((l >> 2) & 1) << B2;
            if (K > 3) idx |= ((l >> 3) & 1) << B3;
            x[l] = v[idx];
        }
#pragma unroll
        for (int r = 0; r < D; r++) {
            cx<R> acc = mk<R>(0, 0);
#pragma unroll
            for (int l = 0; l < D; l++) acc = cfma(acc, m[r * D + l], x[l]);
            int idx = b | ((r & 1) << B0);
            if (K > 1) idx |= ((r >> 1) & 1) << B1;
            if (K > 2) idx |= ((r >> 2) & 1) << B2;
            if (K > 3) idx |= ((r >> 3) & 1) << B3;
            v[idx] = acc;
        }
    }
}

// depolarising / damping on the sweep bits A < B holding (q, q+N) (both maps are
// symmetric in the two roles)
template <typename R, int W, int A, int B>
__device__ __forceinline__ void sw_depol1(cx<R>* v, const TileOp& o) {
    const R c1 = (R)o.f[0], c2 = (R)o.f[1], c3 = (R)o.f[2];
#pragma unroll
    for (int i = 0; i < (1 << W); i++) {
        if (i & ((1 << A) | (1 << B))) continue;
        const int i01 = i | (1 << A), i10 = i | (1 << B), i11 = i01 | (1 << B);
        cx<R> v00 = v[i], v11 = v[i11];
        v[i] = mk<R>(c2 * v00.x + c1 * v11.x, c2 * v00.y + c1 * v11.y);
        v[i01] = cscale(v[i01], c3);
        v[i10] = cscale(v[i10], c3);
        v[i11] = mk<R>(c1 * v00.x + c2 * v11.x, c1 * v00.y + c2 * v11.y);
    }
}

template <typename R, int W, int A, int B>
__device__ __forceinline__ void sw_damp(cx<R>* v, const TileOp& o) {
    const R prob = (R)o.f[0], c1 = (R)o.f[1], c2 = (R)o.f[2];
#pragma unroll
    for (int i = 0; i < (1 << W); i++) {
        if (i & ((1 << A) | (1 << B))) continue;
        const int i01 = i | (1 << A), i10 = i | (1 << B), i11 = i01 | (1 << B);
        cx<R> v00 = v[i], v11 = v[i11];
        v[i] = mk<R>(v00.x + prob * v11.x, v00.y + prob * v11.y);
        v[i01] = cscale(v[i01], c1);
        v[i10] = cscale(v[i10], c1);
        v[i11] = cscale(v11, c2);
    }
}

// pair code A*8 + B for sweep bits A < B < W
#define FSA_PAIRS(W, CALL)                                                                   \
    switch (o.p[0] * 8 + o.p[1]) {                                                          \
    case 1: CALL(0, 1); break;                                                              \
    case 2: if constexpr (W > 2) { CALL(0, 2); } break;                                     \
    case 10: if constexpr (W > 2) { CALL(1, 2); } break;                                    \
    case 3: if constexpr (W > 3) { CALL(0, 3); } break;                                     \
    case 11: if constexpr (W > 3) { CALL(1, 3); } break;                                    \
    case 19: if constexpr (W > 3) { CALL(2, 3); } break;                                    \
    default: break;                                                                         \
    }

template <typename R, int W>
__device__ __forceinline__ void sweep_micro(cx<R>* v, const TileOp& o, uint32_t gb, uint64_t base,
                                            const cx<R>* __restrict__ mats) {
    switch (o.kind) {
    case OP_MAT1: {
        if ((base & o.gmask) != o.gmask) return;
        if ((gb & (uint32_t)o.gq) != (uint32_t)o.gq) return;
        switch (o.p[0]) {
        case 0: sw_mat1<R, W, 0>(v, o); break;
        case 1: sw_mat1<R, W, 1>(v, o); break;
        case 2: if constexpr (W > 2) sw_mat1<R, W, 2>(v, o); break;
        case 3: if constexpr (W > 3) sw_mat1<R, W, 3>(v, o); break;
        default: break;
        }
        return;
    }
    case OP_PAULI: {
        const uint32_t gpar = par64(base & o.gmask) ^ par32(gb & (uint32_t)o.gq);
        switch (o.lmask) {
#define FSA_PC(X) case X: if constexpr (X < (1 << W)) sw_pauli<R, W, X>(v, o, gpar); break;
        FSA_PC(0) FSA_PC(1) FSA_PC(2) FSA_PC(3) FSA_PC(4) FSA_PC(5) FSA_PC(6) FSA_PC(7)
        FSA_PC(8) FSA_PC(9) FSA_PC(10) FSA_PC(11) FSA_PC(12) FSA_PC(13) FSA_PC(14) FSA_PC(15)
#undef FSA_PC
        default: break;
        }
        return;
    }
    case OP_PHASE: {
        const cx<R> f0 = ldc<R>(o.f, 0), f1 = ldc<R>(o.f, 1);
        const uint32_t gpar = par64(base & o.gmask) ^ par32(gb & (uint32_t)o.gq);
#pragma unroll
        for (int i = 0; i < (1 << W); i++) v[i] = cmul(v[i], (gpar ^ par32((uint32_t)i & o.lmask)) ? f1 : f0);
        return;
    }
    case OP_SWAP: {
#define FSA_SW(A, B) sw_swap<R, W, A, B>(v)
        FSA_PAIRS(W, FSA_SW)
#undef FSA_SW
        return;
    }
    case OP_MATK: {
        const cx<R>* m = mats + o.mat_off;
        if (o.k == 2) {
#define FSA_MK2(A, B) sw_matk<R, W, 2, A, B, 0, 0>(v, m)
            FSA_PAIRS(W, FSA_MK2)
#undef FSA_MK2
        } else if (o.k == 3) {
            // the host puts dense gates of at most W - 1 targets into sweeps
            if constexpr (W == 4) {
                switch (o.p[0] * 64 + o.p[1] * 8 + o.p[2]) {
                case 0 * 64 + 1 * 8 + 2: sw_matk<R, W, 3, 0, 1, 2, 0>(v, m); break;
                case 0 * 64 + 1 * 8 + 3: sw_matk<R, W, 3, 0, 1, 3, 0>(v, m); break;
                case 0 * 64 + 2 * 8 + 3: sw_matk<R, W, 3, 0, 2, 3, 0>(v, m); break;
                case 1 * 64 + 2 * 8 + 3: sw_matk<R, W, 3, 1, 2, 3, 0>(v, m); break;
                default: break;
                }
            }
        }
        return;
    }
    case OP_DEPHASE1: {
        const R fac = (R)o.f[0];
#pragma unroll
        for (int i = 0; i < (1 << W); i++)
            if (sw_bit(o.p[0], i, gb, base) ^ sw_bit(o.p[1], i, gb, base)) v[i] = cscale(v[i], fac);
        return;
    }
    case OP_DEPHASE2: {
        const R fac = (R)o.f[0];
#pragma unroll
        for (int i = 0; i < (1 << W); i++) {
            uint32_t f = (sw_bit(o.p[0], i, gb, base) ^ sw_bit(o.p[1], i, gb, base)) |
                         (sw_bit(o.p[2], i, gb, base) ^ sw_bit(o.p[3], i, gb, base));
            if (f) v[i] = cscale(v[i], fac);
        }
        return;
    }
    case OP_DEPOL1: {
#define FSA_DP(A, B) sw_depol1<R, W, A, B>(v, o)
        FSA_PAIRS(W, FSA_DP)
#undef FSA_DP
        return;
    }
    case OP_DAMP: {
#define FSA_DM(A, B) sw_damp<R, W, A, B>(v, o)
        FSA_PAIRS(W, FSA_DM)
#undef FSA_DM
        return;
    }
    default:
        return;
    }
}

template <typename R, int W>
__device__ __forceinline__ void sweep_ops(cx<R>* v, const TileOp& sw, const TileOp* __restrict__ micro, uint32_t gb,
                                          uint64_t base, const cx<R>* __restrict__ mats) {
    const TileOp* mo = micro + sw.mat_off;
    for (uint32_t k = 0; k < sw.k; k++) sweep_micro<R, W>(v, mo[k], gb, base, mats);
}

// tile-local offset of register i of a sweep with bits sb[]
template <int W>
__device__ __forceinline__ uint32_t sw_dep(int i, const uint32_t* sb) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < W; j++) if ((i >> j) & 1) o |= 1u << sb[j];
    return o;
}

// a sweep executed on the tile in LDS (one round trip)
template <typename R, int TB, int W>
__device__ __forceinline__ void sweep_lds(cx<R>* s, uint32_t T, const TileOp& sw, const TileOp* __restrict__ micro,
                                          uint64_t base, const cx<R>* __restrict__ mats) {
    constexpr int G = 1 << W;
    uint32_t sb[W];
    uint32_t smask = 0;
#pragma unroll
    for (int j = 0; j < W; j++) { sb[j] = sw.p[j]; smask |= 1u << sb[j]; }
    for (uint32_t g = threadIdx.x; g < (T >> W); g += TB) {
        const uint32_t gb = ins0_mask32(g, smask);
        cx<R> v[G];
#pragma unroll
        for (int i = 0; i < G; i++) v[i] = s[gb | sw_dep<W>(i, sb)];
        sweep_ops<R, W>(v, sw, micro, gb, base, mats);
#pragma unroll
        for (int i = 0; i < G; i++) s[gb | sw_dep<W>(i, sb)] = v[i];
    }
}

// scatter of a tile-local index to global bit positions
__device__ __forceinline__ uint64_t tile_scatter(uint32_t l, const TileGeom& g) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < g.m; i++) if ((l >> i) & 1u) r |= 1ULL << g.q[i];
    return r;
}

// Persistent, software-pipelined pass.  A workgroup walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...; while the items of tile t run, the loads of the
// next tile are in flight into registers (barriers on gfx950 wait only for LDS
// traffic, so the prefetch is not drained by them).  FIN: the first item is a
// sweep fused with the HBM load (registers in that sweep's layout); FOUT: the
// last item is a sweep fused with the HBM store.  A pass made of a single
// fused sweep never touches LDS.  ROWS = tile size / TB for row-layout loads.
template <typename R, int TB, int ROWS, int W, int VAR, bool SW>
__global__ __launch_bounds__(TB, 4) void k_tile(cx<R>* __restrict__ a, TileGeom geom,
                                                const TileOp* __restrict__ items, unsigned numItems,
                                                const TileOp* __restrict__ micro, const cx<R>* __restrict__ mats,
                                                uint64_t numTiles, uint32_t flags) {
    extern __shared__ unsigned char smem_raw[];
    cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
    constexpr int LOGTB = TB == 1024 ? 10 : TB == 512 ? 9 : TB == 256 ? 8 : 7;
    constexpr int LOGROWS = ROWS >= 16 ? 4 : ROWS >= 8 ? 3 : ROWS >= 4 ? 2 : ROWS >= 2 ? 1 : 0;
    constexpr int G = 1 << W;
    constexpr int NB = ROWS > G ? ROWS : G;
    const uint32_t m = geom.m;
    const uint32_t T = 1u << m;
    const uint32_t tid = threadIdx.x;
    const bool owner = tid < T;
    const bool fin = SW && (flags & 1u) != 0, fout = SW && (flags & 2u) != 0;

    // Addresses are ORs of per-bit values selected by compile-time register
    // indices, so only the per-bit values stay live.
    // row layout: element l = r * TB + tid; row bit j of r is tile bit LOGTB + j
    uint64_t scLo = 0;
    for (uint32_t i = 0; i < m && i < (uint32_t)LOGTB; i++) if ((tid >> i) & 1u) scLo |= 1ULL << geom.q[i];
    uint64_t rowBit[LOGROWS > 0 ? LOGROWS : 1];
#pragma unroll
    for (int j = 0; j < LOGROWS; j++) rowBit[j] = (uint32_t)(LOGTB + j) < m ? 1ULL << geom.q[LOGTB + j] : 0ULL;
    // layout of the first sweep when it is fused with the load
    uint64_t gIn[W];
    uint64_t scIn = 0;
    if (fin) {
        uint32_t mi = 0;
#pragma unroll
        for (int j = 0; j < W; j++) { mi |= 1u << items[0].p[j]; gIn[j] = 1ULL << geom.q[items[0].p[j]]; }
        scIn = tile_scatter(ins0_mask32(tid, mi), geom);
    } else {
#pragma unroll
        for (int j = 0; j < W; j++) gIn[j] = 0;
    }

#define FSA_ROWADDR(r) ([&]() { uint64_t x_ = scLo; _Pragma("unroll") for (int j_ = 0; j_ < LOGROWS; j_++) if (((r) >> j_) & 1) x_ |= rowBit[j_]; return x_; }())
#define FSA_SWADDR(i, g) ([&]() { uint64_t x_ = 0; _Pragma("unroll") for (int j_ = 0; j_ < W; j_++) if (((i) >> j_) & 1) x_ |= g[j_]; return x_; }())
#define FSA_PREFETCH(t)                                                                                 \
    do {                                                                                                \
        const uint64_t pb_ = ins0_mask((t), geom.tileMask);                                             \
        if (fin) {                                                                                      \
            _Pragma("unroll") for (int i = 0; i < G; i++) {                                             \
                cx<R> v_ = a[pb_ | scIn | FSA_SWADDR(i, gIn)]; bx[i] = v_.x; by[i] = v_.y;              \
            }                                                                                           \
        } else {                                                                                        \
            _Pragma("unroll") for (int r = 0; r < ROWS; r++) {                                          \
                cx<R> v_ = a[pb_ | FSA_ROWADDR(r)]; bx[r] = v_.x; by[r] = v_.y;                         \
            }                                                                                           \
        }                                                                                               \
    } while (0)

    // Prefetch loads are unconditional (keeping the buffer in registers):
    // threads beyond a small tile repeat valid addresses of it, and the last
    // prefetch of a workgroup re-reads its current tile.
    R bx[NB], by[NB];
    uint64_t tile = blockIdx.x;
    if (tile >= numTiles) return;
    FSA_PREFETCH(tile);
    for (; tile < numTiles; tile += gridDim.x) {
        const uint64_t base = ins0_mask(tile, geom.tileMask);
        const uint64_t next = tile + gridDim.x < numTiles ? tile + gridDim.x : tile;
        if (!fin) {
            if (owner) {
#pragma unroll
                for (int r = 0; r < ROWS; r++) s[((uint32_t)r << LOGTB) | tid] = mk<R>(bx[r], by[r]);
            }
            __syncthreads();
            FSA_PREFETCH(next);
        }
        for (unsigned it = 0; it < numItems; it++) {
            const TileOp& op = items[it];
            if (!SW || op.kind != OP_SWEEP) {
                tile_apply<R, TB, VAR>(s, T, base, op, mats);
                __syncthreads();
                continue;
            }
            // one code path for every sweep: source = prefetch registers (fused
            // load) or LDS, destination = HBM (fused store) or LDS
            const bool fromRegs = fin && it == 0;
            const bool toHbm = fout && it == numItems - 1;
            uint32_t smask = 0;
#pragma unroll
            for (int j = 0; j < W; j++) smask |= 1u << op.p[j];
            for (uint32_t g = tid; g < (T >> W); g += TB) {
                const uint32_t gb = ins0_mask32(g, smask);
                cx<R> v[G];
                if (fromRegs) {
#pragma unroll
                    for (int i = 0; i < G; i++) v[i] = mk<R>(bx[i], by[i]);
                } else {
#pragma unroll
                    for (int i = 0; i < G; i++) v[i] = s[gb | sw_dep<W>(i, op.p)];
                }
                sweep_ops<R, W>(v, op, micro, gb, base, mats);
                if (toHbm) {
                    const uint64_t sc = base | tile_scatter(gb, geom);
#pragma unroll
                    for (int i = 0; i < G; i++) {
                        uint64_t x = sc;
#pragma unroll
                        for (int j = 0; j < W; j++) if ((i >> j) & 1) x |= 1ULL << geom.q[op.p[j]];
                        a[x] = v[i];
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < G; i++) s[gb | sw_dep<W>(i, op.p)] = v[i];
                }
            }
            if (fromRegs) FSA_PREFETCH(next);
            if (!toHbm) __syncthreads();
        }
        if (!fout && owner) {
#pragma unroll
            for (int r = 0; r < ROWS; r++) a[base | FSA_ROWADDR(r)] = s[((uint32_t)r << LOGTB) | tid];
        }
        __syncthreads();
    }
#undef FSA_ROWADDR
#undef FSA_SWADDR
#undef FSA_PREFETCH
}

}  // namespace

template <typename R>
void launch_tile(hipStream_t s, cx<R>* a, unsigned nq, const TileGeom& geom, const TileOp* ditems,
                 unsigned numItems, const TileOp* dmicro, const cx<R>* dmats, uint32_t flags) {
    constexpr int TB = 512;
    constexpr int W = sweep_bits<R>();
    static int numCU = 0;
    if (numCU == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || numCU <= 0)
            numCU = 256;
    }
    uint64_t numTiles = 1ULL << (nq - geom.m);
    size_t lds = sizeof(cx<R>) << geom.m;
    // enough resident workgroups to fill every CU (LDS allows two 64 KiB tiles per CU)
    uint64_t grid = std::min<uint64_t>(numTiles, (uint64_t)numCU * 2);
    const uint32_t rows = geom.m > 9 ? 1u << (geom.m - 9) : 1u;
    const uint32_t var = (flags >> 2) & 3u;
    const bool sw = (flags & 16u) != 0;
    auto* kern = k_tile<R, TB, 1, W, 0, false>;
#define FSA_PICK(RW)                                                                              \
    kern = var == 2 ? k_tile<R, TB, RW, W, 2, false> : var == 1 ? k_tile<R, TB, RW, W, 1, false>          \
         : sw ? k_tile<R, TB, RW, W, 0, true> : k_tile<R, TB, RW, W, 0, false>;
    switch (rows) {
    case 1: FSA_PICK(1) break;
    case 2: FSA_PICK(2) break;
    case 4: FSA_PICK(4) break;
    case 8: FSA_PICK(8) break;
    default: FSA_PICK((sizeof(R) == 8 ? 8 : 16)) break;
    }
#undef FSA_PICK
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(TB), lds, s, a, geom, ditems, numItems, dmicro, dmats,
                       numTiles, flags);
}

template void launch_tile<double>(hipStream_t, cx<double>*, unsigned, const TileGeom&, const TileOp*, unsigned,
                                  const TileOp*, const cx<double>*, uint32_t);
template void launch_tile<float>(hipStream_t, cx<float>*, unsigned, const TileGeom&, const TileOp*, unsigned,
                                 const TileOp*, const cx<float>*, uint32_t);

}  // namespace fsa
