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
#include "engine.hpp"
#include "launch.hpp"
#include "tile_ops.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>

namespace fsa {

namespace {

using namespace tile;

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
// next tile (into registers) and the write-back of the previous one are in
// flight.  Barriers on gfx950 wait only for LDS traffic, and the items issue
// no vector-memory loads (MFMA matrices are staged in LDS), so nothing waits
// for that traffic before the items end.  SW: the pass has register sweeps
// (OP_SWEEP items, one LDS round trip for a run of ops).  ROWS = tile size / TB.
template <typename R, int TB, int ROWS, int W, int VAR, bool SW>
__global__ __launch_bounds__(TB, (sizeof(R) == 8 && ROWS >= 16) ? 2 : 4) void k_tile(cx<R>* __restrict__ a, TileGeom geom,
                                                const TileOp* __restrict__ items, unsigned numItems,
                                                const TileOp* __restrict__ micro, const cx<R>* __restrict__ mats,
                                                uint64_t numTiles, uint32_t flags) {
    extern __shared__ unsigned char smem_raw[];
    cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
    constexpr int LOGTB = TB == 1024 ? 10 : TB == 512 ? 9 : TB == 256 ? 8 : 7;
    constexpr int LOGROWS = ROWS >= 16 ? 4 : ROWS >= 8 ? 3 : ROWS >= 4 ? 2 : ROWS >= 2 ? 1 : 0;
    const uint32_t m = geom.m;
    const uint32_t T = 1u << m;
    const uint32_t tid = threadIdx.x;
    const bool owner = tid < T;
    // the last item (nDrain = 1) is an OP_DIAG4 applied at the write-back; FP32
    // MFMA kernels (twice the accumulators per thread) have no registers for it
    constexpr bool DRAIN = !(sizeof(R) == 4 && VAR >= 2);
    const unsigned nDrain = DRAIN ? (flags >> 8) & 0xffu : 0u;
    // flags bit 6: the tiles are walked in bit-reversed index order, so the
    // tiles in flight at once differ in the highest qubits outside the tile
    // rather than the lowest (the HBM access pattern that measured faster).
    // FP64 only: in the FP32 plain kernel the same few scalar instructions per
    // tile cost 13% (24-qubit passes, 335 against 297 us per launch) with
    // identical vector code, registers and occupancy, so FP32 ignores the bit
    const bool revOrder = (flags & 64u) != 0;
    const uint32_t tileBits = 63u - (uint32_t)__builtin_clzll(numTiles | 1ULL);
    auto tile_index = [&](uint64_t t) -> uint64_t {
        if constexpr (sizeof(R) == 8) {
            if (revOrder && tileBits) return (uint64_t)__builtin_bitreverse64(t) >> (64u - tileBits);
        }
        return t;
    };
    const unsigned nBody = numItems - nDrain;

    // Addresses are ORs of per-bit values selected by compile-time register
    // indices, so only the per-bit values stay live.
    // row layout: element l = r * TB + tid; row bit j of r is tile bit LOGTB + j
    uint64_t scLo = 0;
    for (uint32_t i = 0; i < m && i < (uint32_t)LOGTB; i++) if ((tid >> i) & 1u) scLo |= 1ULL << geom.q[i];
    uint64_t rowBit[LOGROWS > 0 ? LOGROWS : 1];
#pragma unroll
    for (int j = 0; j < LOGROWS; j++) rowBit[j] = (uint32_t)(LOGTB + j) < m ? 1ULL << geom.q[LOGTB + j] : 0ULL;

#define FSA_ROWADDR(r) ([&]() { uint64_t x_ = scLo; _Pragma("unroll") for (int j_ = 0; j_ < LOGROWS; j_++) if (((r) >> j_) & 1) x_ |= rowBit[j_]; return x_; }())
#define FSA_PREFETCH(t)                                                                                 \
    do {                                                                                                \
        const uint64_t pb_ = ins0_mask(tile_index(t), geom.tileMask);                                   \
        _Pragma("unroll") for (int r = 0; r < ROWS; r++) {                                              \
            cx<R> v_ = a[pb_ | FSA_ROWADDR(r)]; bx[r] = v_.x; by[r] = v_.y;                             \
        }                                                                                               \
    } while (0)

    // Threads beyond a small tile load valid addresses of it (repeats) and
    // neither store nor write LDS.
    R bx[ROWS], by[ROWS];
    uint64_t tile = blockIdx.x;
    if (tile >= numTiles) return;
    // MFMA matrices go to LDS behind the tile, at their item's offset, in the
    // order mat_lds_index reads them: the left half of the real form, or for
    // the FP64 three-product form (tile_mfma3_form) U itself as complex
    // entries.  VAR 2 copies every matrix of the pass once; VAR 3 (matrices
    // beyond the LDS area) copies each one just before its item, at offset 0.
    R* lmat = reinterpret_cast<R*>(smem_raw + ((size_t)sizeof(cx<R>) << m));
    auto stage_matrix = [&](const TileOp& op) {
        const uint32_t k = op.k, D = 1u << k;
        const R* M = reinterpret_cast<const R*>(mats + op.mat_off);
        if (tile_mfma3_form<R, TB>(k)) {
            for (uint32_t e = tid; e < D * D; e += TB) {
                const uint32_t r = e >> k, c = e & (D - 1);
                R* dst = lmat + op.lmask + 2 * (r * D + (c ^ (r & (D - 1))));
                dst[0] = M[(size_t)r * 2 * D + c];
                dst[1] = M[(size_t)(r + D) * 2 * D + c];
            }
            return;
        }
        for (uint32_t e = tid; e < 2 * D * D; e += TB) {
            const uint32_t r = e >> k, c = e & (D - 1);
            lmat[op.lmask + r * D + (c ^ (r & (D - 1)))] = M[(size_t)r * 2 * D + c];
        }
    };
    if constexpr (VAR == 2) {
        for (unsigned it = 0; it < numItems; it++) {
            const TileOp& op = items[it];
            if (op.kind == OP_MATK && op.lmask2 == 1u && op.lmask != MAT_UNPLACED) stage_matrix(op);
        }
        __syncthreads();
    }
    FSA_PREFETCH(tile);
    if (owner) {
#pragma unroll
        for (int r = 0; r < ROWS; r++) s[((uint32_t)r << LOGTB) | tid] = mk<R>(bx[r], by[r]);
    }
    __syncthreads();
    uint64_t next = tile + gridDim.x;
    if (next < numTiles) FSA_PREFETCH(next);
    for (;;) {
        const uint64_t base = ins0_mask(tile_index(tile), geom.tileMask);
        for (unsigned it = 0; it < nBody; it++) {
            const TileOp& op = items[it];
            if constexpr (SW) {
                if (op.kind == OP_SWEEP) {
                    sweep_lds<R, TB, W>(s, T, op, micro, base, mats);
                    __syncthreads();
                    continue;
                }
            }
            if constexpr (VAR == 3) {
                // the previous item's barrier: every wave is done with the area
                if (op.kind == OP_MATK && op.lmask2 == 1u) {
                    stage_matrix(op);
                    __syncthreads();
                }
            }
            // batched simple items, as many pairs per batch as the kernel has
            // registers for (none in the MFMA kernels)
            constexpr int GP = (TB != 512 || SW || VAR >= 2) ? 0
                             : sizeof(R) == 4 ? (VAR == 0 ? 1 : 0)
                             : VAR == 0 ? 4 : 2;
            tile_apply<R, TB, VAR, (ROWS >= 16 ? ROWS / 8 : 1), R, (GP ? ROWS : 0), GP>(s, T, base, op, mats, lmat);
            __syncthreads();
        }
        // The write-back of this tile and the loads of the tile after the next
        // are issued together, so both complete while the next tile's items
        // run.  (Vector memory operations complete in order: stores issued
        // just before the next tile is moved into LDS would be waited for
        // there, leaving their latency exposed on every tile.)
        // Each thread moves its own LDS slots out (to HBM) and the next tile in
        // (from its registers), so no barrier separates the two; the wait below
        // makes the next tile's loads land before the first store of this tile
        // is issued.
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        const bool more = next < numTiles;
        if (owner && nDrain == 0 && more) {
#pragma unroll
            for (int r = 0; r < ROWS; r++) {
                const uint32_t l = ((uint32_t)r << LOGTB) | tid;
                const R vx = s[l].x, vy = s[l].y;
                s[l] = mk<R>(bx[r], by[r]);
                a[base | FSA_ROWADDR(r)] = mk<R>(vx, vy);
            }
        } else if (owner && nDrain == 0) {
#pragma unroll
            for (int r = 0; r < ROWS; r++) a[base | FSA_ROWADDR(r)] = s[((uint32_t)r << LOGTB) | tid];
        } else if (owner) {
            // the pass-end OP_DIAG4 item: amplitude l is multiplied by
            // f[(p1 << 1) | p2], p1 / p2 the parities under its two masks
            const TileOp& dop = items[nBody];
            const uint32_t m1 = dop.lmask, m2 = dop.lmask2;
            const uint32_t b1 = par64(base & dop.gmask), b2 = par64(base & dop.gq);
            R fx0 = (R)dop.f[0], fy0 = (R)dop.f[1], fx1 = (R)dop.f[2], fy1 = (R)dop.f[3];
            R fx2 = (R)dop.f[4], fy2 = (R)dop.f[5], fx3 = (R)dop.f[6], fy3 = (R)dop.f[7];
            // empty register constraints (no instruction): without them the
            // compiler turns the per-amplitude selects among these loaded
            // values into an indexed load from a private copy of the table
            asm volatile("" : "+v"(fx0), "+v"(fy0), "+v"(fx1), "+v"(fy1));
            asm volatile("" : "+v"(fx2), "+v"(fy2), "+v"(fx3), "+v"(fy3));
#pragma unroll
            for (int r = 0; r < ROWS; r++) {
                const uint32_t l = ((uint32_t)r << LOGTB) | tid;
                const R vx = s[l].x, vy = s[l].y;
                if (more) s[l] = mk<R>(bx[r], by[r]);
                const uint32_t p1 = par32(l & m1) ^ b1, p2 = par32(l & m2) ^ b2;
                const R fx = p1 ? (p2 ? fx3 : fx2) : (p2 ? fx1 : fx0);
                const R fy = p1 ? (p2 ? fy3 : fy2) : (p2 ? fy1 : fy0);
                a[base | FSA_ROWADDR(r)] = cmul(mk<R>(vx, vy), mk<R>(fx, fy));
            }
        }
        if (!more) break;
        if (next + gridDim.x < numTiles) FSA_PREFETCH(next + gridDim.x);
        __syncthreads();   // the next tile is in LDS
        tile = next;
        next += gridDim.x;
    }
#undef FSA_ROWADDR
#undef FSA_PREFETCH
}

// Resident workgroups per CU for a tile kernel at this LDS size (cached; at most
// 4).  FSA_TILE_OCC=0 keeps the fixed count `dflt` (two 512-thread workgroups,
// or one of 1024).
unsigned tile_workgroups_per_cu(const void* kern, unsigned tb, size_t lds, unsigned dflt) {
    const char* e = std::getenv("FSA_TILE_OCC");
    if (e && e[0] == '0') return dflt;
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, unsigned> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kern, lds});
    if (it != cache.end()) return it->second;
    int occ = 0;
    unsigned n = dflt;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, (int)tb, lds) == hipSuccess && occ > 0)
        n = std::min(4u, (unsigned)occ);
    cache[{kern, lds}] = n;
    return n;
}

// kernel of a pass: TB threads, ROWS = tile / TB amplitudes per thread
template <typename R, int TB, int RW>
auto pick_tile(uint32_t var, bool sw) {
    constexpr int W = sweep_bits<R>();
    return var == 3 ? k_tile<R, TB, RW, W, 3, false> : var == 2 ? k_tile<R, TB, RW, W, 2, false>
         : var == 1 ? k_tile<R, TB, RW, W, 1, false>
         : sw ? k_tile<R, TB, RW, W, 0, true> : k_tile<R, TB, RW, W, 0, false>;
}

}  // namespace

template <typename R>
void launch_tile(hipStream_t s, cx<R>* a, unsigned nq, const TileGeom& geom, const TileOp* ditems,
                 unsigned numItems, const TileOp* dmicro, const cx<R>* dmats, uint32_t flags, uint32_t matLds) {
    static int numCU = 0;
    if (numCU == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || numCU <= 0)
            numCU = 256;
    }
    uint64_t numTiles = 1ULL << (nq - geom.m);
    const uint32_t var = (flags >> 2) & 3u;
    const size_t tileBytes = sizeof(cx<R>) << geom.m;
    if (matLds > (var >= 2 ? tile_mat_lds_bytes(tileBytes) : 0u))
        fail(FSA_ERR_INTERNAL, "tile pass asks for more LDS for matrices than its variant has");
    const size_t lds = tileBytes + matLds;
    // 512-thread workgroups for tiles of up to 64 KiB (two per CU at 64 KiB), one
    // of 1024 threads for a 128 KiB tile
    const bool big = tileBytes > 64 * 1024;
    const uint32_t tb = big ? 1024u : 512u;
    const bool sw = (flags & 16u) != 0;
    // FP64 tiles in bit-reversed order unless FSA_TILE_ORDER=0 (30-qubit 1000-gate
    // circuit: 791-793 against 774 gates/s on one box, 777 against 760 on another);
    // the FP32 kernels always walk the index order
    const char* orderEnv = std::getenv("FSA_TILE_ORDER");
    if (!(orderEnv && orderEnv[0] == '0')) flags |= 64u;
    auto* kern = k_tile<R, 512, 1, sweep_bits<R>(), 0, false>;
    if (big) {
        if constexpr (sizeof(R) == 8) {
            if (geom.m != 13) fail(FSA_ERR_INTERNAL, "tile of more than 64 KiB that is not 2^13 FP64 amplitudes");
            kern = pick_tile<R, 1024, 8>(var, sw);
        } else {
            fail(FSA_ERR_INTERNAL, "FP32 tiles hold at most 64 KiB");
        }
    } else {
        switch (geom.m > 9 ? 1u << (geom.m - 9) : 1u) {
        case 1: kern = pick_tile<R, 512, 1>(var, sw); break;
        case 2: kern = pick_tile<R, 512, 2>(var, sw); break;
        case 4: kern = pick_tile<R, 512, 4>(var, sw); break;
        case 8: kern = pick_tile<R, 512, 8>(var, sw); break;
        default: kern = pick_tile<R, 512, 16>(var, sw); break;
        }
    }
    if (lds > 64 * 1024) {
        // dynamic LDS beyond 64 KiB must be granted per kernel before the launch
        static std::mutex mu;
        static std::unordered_map<const void*, size_t> granted;
        std::lock_guard<std::mutex> lock(mu);
        size_t& g = granted[reinterpret_cast<const void*>(kern)];
        if (lds > g) {
            if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                fail(FSA_ERR_HIP, "cannot grant the tile kernel more than 64 KiB of LDS");
            g = lds;
        }
    }
    // The persistent grid is as many workgroups as fit on the GPU at once: what
    // the kernel's registers and this pass's LDS allow per CU (smaller tiles
    // leave room for more), so every workgroup is resident from the start.
    const unsigned perCU = tile_workgroups_per_cu(reinterpret_cast<const void*>(kern), tb, lds, big ? 1u : 2u);
    const uint64_t grid = std::min<uint64_t>(numTiles, (uint64_t)numCU * perCU);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(tb), lds, s, a, geom, ditems, numItems, dmicro, dmats,
                       numTiles, flags);
}

template void launch_tile<double>(hipStream_t, cx<double>*, unsigned, const TileGeom&, const TileOp*, unsigned,
                                  const TileOp*, const cx<double>*, uint32_t, uint32_t);
template void launch_tile<float>(hipStream_t, cx<float>*, unsigned, const TileGeom&, const TileOp*, unsigned,
                                 const TileOp*, const cx<float>*, uint32_t, uint32_t);

}  // namespace fsa
