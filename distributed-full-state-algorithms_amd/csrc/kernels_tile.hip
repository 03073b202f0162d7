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

namespace fsa {

namespace {

// bit value of a channel qubit: code bit 8 set -> local position in the tile,
// otherwise a global qubit index outside the tile (fixed for the workgroup)
__device__ __forceinline__ uint32_t chan_bit(uint32_t l, uint64_t base, uint32_t code) {
    return (code & 0x100u) ? (l >> (code & 0xffu)) & 1u : (uint32_t)((base >> code) & 1ULL);
}

template <typename R>
__device__ __forceinline__ cx<R> ldc(const double* f, int i) { return mk<R>((R)f[2 * i], (R)f[2 * i + 1]); }

template <typename R, int K>
__device__ void tile_matk(cx<R>* s, uint32_t T, const TileOp& op, const cx<R>* __restrict__ mat) {
    constexpr int D = 1 << K;
    uint32_t tm = 0;
    uint32_t off[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < K; i++) if ((l >> i) & 1) o |= 1u << op.p[i];
        off[l] = o;
    }
#pragma unroll
    for (int i = 0; i < K; i++) tm |= 1u << op.p[i];
    for (uint32_t g = threadIdx.x; g < (T >> K); g += BLOCK) {
        uint32_t b = ins0_mask32(g, tm);
        cx<R> v[D];
#pragma unroll
        for (int l = 0; l < D; l++) v[l] = s[b | off[l]];
#pragma unroll
        for (int r = 0; r < D; r++) {
            cx<R> acc = mk<R>(0, 0);
#pragma unroll
            for (int l = 0; l < D; l++) acc = cfma(acc, mat[r * D + l], v[l]);
            s[b | off[r]] = acc;
        }
    }
}

template <typename R>
__device__ void tile_apply(cx<R>* s, uint32_t T, uint64_t base, const TileOp& op,
                           const cx<R>* __restrict__ mats) {
    const uint32_t tid = threadIdx.x;
    switch (op.kind) {
    case OP_MAT1: {
        if ((base & op.gmask) != op.gmask) return;           // an outside control is 0
        const uint32_t tb = 1u << op.p[0];
        const uint32_t lc = op.lmask;
        const uint32_t nIts = T >> (1 + __builtin_popcount(lc));
        const cx<R> g00 = ldc<R>(op.f, 0), g01 = ldc<R>(op.f, 1), g10 = ldc<R>(op.f, 2), g11 = ldc<R>(op.f, 3);
        for (uint32_t j = tid; j < nIts; j += BLOCK) {
            uint32_t i0 = ins0_mask32(j, lc | tb) | lc;
            cx<R> v0 = s[i0], v1 = s[i0 | tb];
            s[i0] = cfma(cmul(g00, v0), g01, v1);
            s[i0 | tb] = cfma(cmul(g10, v0), g11, v1);
        }
        return;
    }
    case OP_SWAP: {
        const uint32_t b0 = 1u << op.p[0], b1 = 1u << op.p[1];
        for (uint32_t j = tid; j < (T >> 2); j += BLOCK) {
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
            for (uint32_t l = tid; l < T; l += BLOCK) {
                cx<R> v = s[l];
                cx<R> f = (bp ^ par32(l & lyz)) ? fo : fe;
                s[l] = cfma(cmul(th, v), f, v);
            }
            return;
        }
        const unsigned hb = 31u - (unsigned)__builtin_clz(lxy);
        for (uint32_t j = tid; j < (T >> 1); j += BLOCK) {
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
        for (uint32_t l = tid; l < T; l += BLOCK)
            s[l] = cmul(s[l], (bp ^ par32(l & op.lmask)) ? f1 : f0);
        return;
    }
    case OP_MATK: {
        const cx<R>* mat = mats + op.mat_off;
        switch (op.k) {
        case 1: tile_matk<R, 1>(s, T, op, mat); break;
        case 2: tile_matk<R, 2>(s, T, op, mat); break;
        case 3: tile_matk<R, 3>(s, T, op, mat); break;
        case 4: tile_matk<R, 4>(s, T, op, mat); break;
        case 5: tile_matk<R, 5>(s, T, op, mat); break;
        default: break;
        }
        return;
    }
    case OP_DEPHASE1: {
        const R fac = (R)op.f[0];
        const uint32_t cq = op.p[0], ca = op.p[1];
        for (uint32_t l = tid; l < T; l += BLOCK)
            if (chan_bit(l, base, cq) ^ chan_bit(l, base, ca)) s[l] = cscale(s[l], fac);
        return;
    }
    case OP_DEPHASE2: {
        const R fac = (R)op.f[0];
        for (uint32_t l = tid; l < T; l += BLOCK) {
            uint32_t f = (chan_bit(l, base, op.p[0]) ^ chan_bit(l, base, op.p[1])) |
                         (chan_bit(l, base, op.p[2]) ^ chan_bit(l, base, op.p[3]));
            if (f) s[l] = cscale(s[l], fac);
        }
        return;
    }
    case OP_DEPOL1: {
        const R c1 = (R)op.f[0], c2 = (R)op.f[1], c3 = (R)op.f[2];
        const uint32_t bq = 1u << op.p[0], ba = 1u << op.p[1];
        for (uint32_t j = tid; j < (T >> 2); j += BLOCK) {
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
        for (uint32_t j = tid; j < (T >> 2); j += BLOCK) {
            uint32_t i = ins0_mask32(j, bq | ba);
            cx<R> v00 = s[i], v11 = s[i | bq | ba];
            s[i] = mk<R>(v00.x + prob * v11.x, v00.y + prob * v11.y);
            s[i | bq] = cscale(s[i | bq], c1);
            s[i | ba] = cscale(s[i | ba], c1);
            s[i | bq | ba] = cscale(v11, c2);
        }
        return;
    }
    case OP_DEPOL2: {
        const R c1 = (R)op.f[0], c2 = (R)op.f[1], sc = (R)op.f[2];
        const uint32_t b[4] = {1u << op.p[0], 1u << op.p[1], 1u << op.p[2], 1u << op.p[3]};
        const uint32_t mk4 = b[0] | b[1] | b[2] | b[3];
        for (uint32_t j = tid; j < (T >> 4); j += BLOCK) {
            uint32_t i = ins0_mask32(j, mk4);
            cx<R> v[16];
#pragma unroll
            for (int l = 0; l < 16; l++) {
                uint32_t x = i;
#pragma unroll
                for (int q = 0; q < 4; q++) if ((l >> q) & 1) x |= b[q];
                v[l] = s[x];
            }
            cx<R> sum = cadd(cadd(v[0], v[5]), cadd(v[10], v[15]));
#pragma unroll
            for (int l = 0; l < 16; l++) {
                uint32_t x = i;
#pragma unroll
                for (int q = 0; q < 4; q++) if ((l >> q) & 1) x |= b[q];
                bool diag = (l == 0 || l == 5 || l == 10 || l == 15);
                s[x] = diag ? mk<R>(c1 * v[l].x + c2 * sum.x, c1 * v[l].y + c2 * sum.y) : cscale(v[l], sc);
            }
        }
        return;
    }
    default:
        return;
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_tile(cx<R>* __restrict__ a, TileGeom geom,
                                                const TileOp* __restrict__ ops, unsigned numOps,
                                                const cx<R>* __restrict__ mats) {
    extern __shared__ unsigned char smem_raw[];
    cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
    const uint32_t m = geom.m;
    const uint32_t T = 1u << m;
    const uint32_t tid = threadIdx.x;
    const uint64_t base = ins0_mask((uint64_t)blockIdx.x, geom.tileMask);

    // scatter of the low 8 tile-index bits (thread-constant) and of the rest (uniform)
    uint64_t scLo = 0;
    for (uint32_t i = 0; i < m && i < 8; i++) if ((tid >> i) & 1u) scLo |= 1ULL << geom.q[i];
    const uint32_t rows = T > BLOCK ? T / BLOCK : 1;

#pragma unroll 16
    for (uint32_t r = 0; r < rows; r++) {
        uint64_t scHi = 0;
        for (uint32_t i = 8; i < m; i++) if ((r >> (i - 8)) & 1u) scHi |= 1ULL << geom.q[i];
        uint32_t l = (r << 8) | tid;
        if (l < T) s[l] = a[base | scLo | scHi];
    }
    for (unsigned o = 0; o < numOps; o++) {
        __syncthreads();
        tile_apply<R>(s, T, base, ops[o], mats);
    }
    __syncthreads();
#pragma unroll 16
    for (uint32_t r = 0; r < rows; r++) {
        uint64_t scHi = 0;
        for (uint32_t i = 8; i < m; i++) if ((r >> (i - 8)) & 1u) scHi |= 1ULL << geom.q[i];
        uint32_t l = (r << 8) | tid;
        if (l < T) a[base | scLo | scHi] = s[l];
    }
}

}  // namespace

template <typename R>
void launch_tile(hipStream_t s, cx<R>* a, unsigned nq, const TileGeom& geom, const TileOp* dops,
                 unsigned numOps, const cx<R>* dmats) {
    uint64_t numTiles = 1ULL << (nq - geom.m);
    size_t lds = sizeof(cx<R>) << geom.m;
    hipLaunchKernelGGL(k_tile<R>, dim3((unsigned)numTiles), dim3(BLOCK), lds, s, a, geom, dops, numOps, dmats);
}

template void launch_tile<double>(hipStream_t, cx<double>*, unsigned, const TileGeom&, const TileOp*,
                                  unsigned, const cx<double>*);
template void launch_tile<float>(hipStream_t, cx<float>*, unsigned, const TileGeom&, const TileOp*,
                                 unsigned, const cx<float>*);

}  // namespace fsa
