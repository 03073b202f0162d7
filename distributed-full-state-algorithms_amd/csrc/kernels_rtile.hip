// kernels_rtile.hip -- fused pass over tiles held in registers (FP64).
//
// A tile of 2^12 amplitudes (tile-local index l spells the pass's 12 tile
// qubits, as in kernels_tile.hip) is spread over the 512 threads of a
// workgroup, 8 amplitudes per thread, and stays in registers from its load to
// its store.  Where an amplitude sits is its slot s = lane | wave << 6 | r << 9
// (r = the register); a layout says which tile bit each of the 12 slot bits
// holds (4 bits per slot, packed in a 64-bit word).  An operation whose mixed
// qubits all sit on lane or register slots runs without LDS: register slots
// pair registers of one thread, lane slots pair lanes of one wave through
// cross-lane exchanges, and diagonal factors, controls and Z signs read the
// slot bits directly.  When an operation needs a qubit that sits on a wave
// slot, the host puts an OP_REMAP item in front of it: the tile goes through
// LDS once and comes back in a layout with that qubit local, so a pass costs
// one LDS round trip per remap instead of one per operation.
//
// Dense gates of 2-5 targets run on MFMA (v_mfma_f64_16x16x4f64) with the B
// operands read straight from the registers: the host lays matrix bits 0 and 1
// on lane slots 4 and 5 (a B operand's row within a k-step is lane / 16) and
// matrix bits 2.. on register slots 9.. (the k-step), so the accumulators
// (row 4j + lane / 16 of each 16-row block) come back in the same layout.  A
// wave multiplies the real form [[Ur, -Ui], [Ui, Ur]] (in LDS, staged once per
// workgroup) by its 16-column blocks: lane % 16 and the register bits that
// hold no matrix bit pick the columns.
#include "engine.hpp"
#include "launch.hpp"
#include "tile_ops.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

namespace fsa {

namespace {

using namespace tile;

constexpr int RT_G = 8;                       // amplitudes per thread
constexpr uint32_t RT_REG0 = 9;               // first register slot (6 lane + 3 wave slots below)
constexpr uint32_t RT_TILE_BYTES = 65536;     // one remap buffer half: a tile of 2^12 FP64 amplitudes

__device__ __forceinline__ uint32_t rt_nib(uint64_t map, uint32_t s) { return (uint32_t)(map >> (4 * s)) & 15u; }

__device__ __forceinline__ cx<double> rt_shx(cx<double> v, uint32_t m) {
    return mk<double>(__shfl_xor(v.x, (int)m, 64), __shfl_xor(v.y, (int)m, 64));
}

// component-wise select: a select between two complex structs would compile to
// a select of their addresses, which keeps the register arrays out of registers
__device__ __forceinline__ cx<double> rt_sel(bool c, cx<double> a, cx<double> b) {
    return mk<double>(c ? a.x : b.x, c ? a.y : b.y);
}

// register r's part of a slot index / of a tile index given per-register-slot values
__device__ __forceinline__ uint32_t rt_rslot(int r) { return (uint32_t)r << RT_REG0; }

// one-target gate on register slot RT_REG0 + J (pairs registers of one thread)
template <int J>
__device__ __forceinline__ void rt_mat1_reg(cx<double> (&v)[RT_G], const TileOp& op, uint32_t lw) {
    const cx<double> g00 = ldc<double>(op.f, 0), g01 = ldc<double>(op.f, 1);
    const cx<double> g10 = ldc<double>(op.f, 2), g11 = ldc<double>(op.f, 3);
    const uint32_t cm = op.lmask;
    const uint32_t cl = cm & ((1u << RT_REG0) - 1u), cr = cm >> RT_REG0;   // thread and register parts
    const bool okl = (lw & cl) == cl;
#pragma unroll
    for (int r = 0; r < RT_G; r++) {
        if (r & (1 << J)) continue;
        if (((uint32_t)r & cr) != cr) continue;              // a register control is 0 (uniform)
        const int r1 = r | (1 << J);
        const cx<double> a0 = v[r], a1 = v[r1];
        const cx<double> b0 = cfma(cmul(g00, a0), g01, a1), b1 = cfma(cmul(g10, a0), g11, a1);
        if (cl) {                                            // controls never include the target slot
            v[r] = rt_sel(okl, b0, a0);
            v[r1] = rt_sel(okl, b1, a1);
        } else {
            v[r] = b0;
            v[r1] = b1;
        }
    }
}

// one-target gate on lane slot ts (< 6): the partner amplitude is in lane ^ 2^ts
__device__ __forceinline__ void rt_mat1_lane(cx<double> (&v)[RT_G], const TileOp& op, uint32_t lw, uint32_t ts) {
    const uint32_t lm = 1u << ts;
    const bool hi = ((lw >> ts) & 1u) != 0;
    // this lane computes row `hi` of the 2x2: own value times U[hi][hi], partner's times U[hi][!hi]
    const cx<double> gs = rt_sel(hi, ldc<double>(op.f, 3), ldc<double>(op.f, 0));
    const cx<double> gp = rt_sel(hi, ldc<double>(op.f, 2), ldc<double>(op.f, 1));
    const uint32_t cm = op.lmask;
    const uint32_t cl = cm & ((1u << RT_REG0) - 1u), cr = cm >> RT_REG0;
    const bool okl = (lw & cl) == cl;
#pragma unroll
    for (int r = 0; r < RT_G; r++) {
        if (((uint32_t)r & cr) != cr) continue;              // a register control is 0 (uniform)
        const cx<double> p = rt_shx(v[r], lm);
        const cx<double> nv = cfma(cmul(gs, v[r]), gp, p);
        v[r] = cl ? rt_sel(okl, nv, v[r]) : nv;
    }
}

// Pauli tensor / gadget with X/Y mask mxy (slots): new v(s) = th v(s) + f(s ^ mxy) v(s ^ mxy),
// f(k) = -fe or fe by the parity of k & myz (and of the tile base under the
// out-of-tile mask); RX = the mask's register part, its lane part is exchanged
template <int RX>
__device__ __forceinline__ void rt_pauli(cx<double> (&v)[RT_G], const TileOp& op, uint32_t lw, uint32_t bp) {
    const cx<double> th = ldc<double>(op.f, 0), fe = ldc<double>(op.f, 1);
    const cx<double> fo = mk<double>(-fe.x, -fe.y);
    const uint32_t lx = op.lmask & 63u, myz = op.lmask2;
    // the partner's parity: its lane/wave part (per thread) and its register part (uniform)
    const uint32_t pp = bp ^ par32((lw ^ lx) & myz & ((1u << RT_REG0) - 1u));
    const uint32_t myzr = myz >> RT_REG0;
#pragma unroll
    for (int r = 0; r < RT_G; r++) {
        const int r2 = r ^ RX;
        if (r2 < r) continue;
        if (RX == 0) {
            const cx<double> p = lx ? rt_shx(v[r], lx) : v[r];   // lx is uniform
            const cx<double> f = rt_sel((pp ^ par32((uint32_t)r & myzr)) != 0, fo, fe);
            v[r] = cfma(cmul(th, v[r]), f, p);
        } else {
            // (lane, r) pairs with (lane ^ lx, r2) and (lane, r2) with (lane ^ lx, r)
            const cx<double> p = lx ? rt_shx(v[r2], lx) : v[r2];
            const cx<double> q = lx ? rt_shx(v[r], lx) : v[r];
            const cx<double> f = rt_sel((pp ^ par32((uint32_t)r2 & myzr)) != 0, fo, fe);
            const cx<double> g = rt_sel((pp ^ par32((uint32_t)r & myzr)) != 0, fo, fe);
            const cx<double> a = v[r], b = v[r2];
            v[r] = cfma(cmul(th, a), f, p);
            v[r2] = cfma(cmul(th, b), g, q);
        }
    }
}

__device__ __forceinline__ void rt_phase(cx<double> (&v)[RT_G], const TileOp& op, uint32_t lw, uint32_t bp) {
    const cx<double> f0 = ldc<double>(op.f, 0), f1 = ldc<double>(op.f, 1);
    const uint32_t pl = bp ^ par32(lw & op.lmask & ((1u << RT_REG0) - 1u)), mr = op.lmask >> RT_REG0;
    const cx<double> e0 = rt_sel(pl != 0, f1, f0), e1 = rt_sel(pl != 0, f0, f1);   // by the register part's parity
#pragma unroll
    for (int r = 0; r < RT_G; r++) v[r] = cmul(v[r], par32((uint32_t)r & mr) ? e1 : e0);
}

// dense K-target gate on MFMA: matrix bits 0, 1 on lane slots 4, 5, bits 2..K-1
// on register slots 9.. (host layout); M = the left half [Ur; Ui] of the real
// form in LDS (mat_lds_index order)
template <int K>
__device__ __forceinline__ void rt_mfma(cx<double> (&v)[RT_G], const double* __restrict__ M) {
    constexpr int D = 1 << K, D2 = 2 * D;
    constexpr int RB = D2 >= 16 ? D2 / 16 : 1;          // 16-row blocks (K = 2: one, rows 8-15 empty)
    constexpr int KS = D2 / 4;                           // k-steps over the real rows
    constexpr int HALF = KS / 2;                         // k-steps of the real parts (D / 4)
    constexpr int TR = K - 2;                            // register bits holding matrix bits 2..K-1
    constexpr int NCB = RT_G >> TR;                      // column blocks: values of the other register bits
    typedef MfmaAcc<double>::T acc_t;
    const uint32_t lane = threadIdx.x & 63u, lr = lane & 15u, lk = lane >> 4;
    // A[row][col] of [[Ur, -Ui], [Ui, Ur]] from the left half: col < D -> left[row][col];
    // col >= D -> -left[row + D][col - D] (row < D) or left[row - D][col - D]
    auto a_at = [&](uint32_t row, uint32_t col) -> double {
        if (row >= (uint32_t)D2) return 0.0;
        if (col < (uint32_t)D) return M[mat_lds_index<D>(row, col)];
        return row < (uint32_t)D ? -M[mat_lds_index<D>(row + D, col - D)] : M[mat_lds_index<D>(row - D, col - D)];
    };
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
        // row blocks one at a time (its A operands only are in flight); the
        // outputs wait in acc until every row block has read the inputs
        acc_t acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            acc[rb] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const int reg = (ks % HALF) | (cb << TR);
                const double b = ks < HALF ? v[reg].x : v[reg].y;
                acc[rb] = MfmaAcc<double>::step(a_at((uint32_t)rb * 16 + lr, (uint32_t)ks * 4 + lk), b, acc[rb]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // accumulator j of lane l is row 16 rb + 4 j + l / 16: real rows < D, imaginary rows D..2D-1
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row4 = 16 * rb + 4 * j;
                if (row4 < D) v[(row4 >> 2) | (cb << TR)].x = acc[rb][j];
                else if (row4 < D2) v[((row4 - D) >> 2) | (cb << TR)].y = acc[rb][j];
            }
    }
}

// new layout: every amplitude goes to LDS at its old slot index (each wave
// writes 64 contiguous elements per register) and comes back from the old
// slot of the tile bit each new slot holds (op.gmask: old slot of new slot s,
// 4 bits per slot; a remap usually moves few bits, so reads stay mostly
// contiguous too).  Consecutive remaps use the two halves of buf in turn, so
// one barrier per remap suffices: a half is written again only after every
// thread has passed the next remap's barrier, that is after its reads.
__device__ __forceinline__ void rt_remap(cx<double> (&v)[RT_G], const TileOp& op, uint32_t lw, cx<double>* buf) {
    const uint64_t perm = op.gmask;
    uint32_t src = 0;
#pragma unroll
    for (uint32_t b = 0; b < RT_REG0; b++) src |= ((lw >> b) & 1u) << rt_nib(perm, b);
    const uint32_t p0 = 1u << rt_nib(perm, RT_REG0), p1 = 1u << rt_nib(perm, RT_REG0 + 1),
                   p2 = 1u << rt_nib(perm, RT_REG0 + 2);
#pragma unroll
    for (int r = 0; r < RT_G; r++) buf[lw | rt_rslot(r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RT_G; r++) v[r] = buf[src | ((r & 1) ? p0 : 0u) | ((r & 2) ? p1 : 0u) | ((r & 4) ? p2 : 0u)];
}

// Persistent pass.  PF: software-pipelined, one workgroup per CU: while the
// items of tile t run in registers, the loads of tile t + 1 (into nx) and the
// stores of tile t - 1 are in flight.  Without PF each tile is loaded when it
// starts and two workgroups share a CU (at most 128 VGPRs), so one's memory
// traffic overlaps the other's items.  loadMap / storeMap: the layouts of the
// load and of the store.
template <bool PF>
__global__ __launch_bounds__(512, PF ? 1 : 4) void k_rtile(cx<double>* __restrict__ a, TileGeom geom,
                                                  const TileOp* __restrict__ items, unsigned numItems,
                                                  const cx<double>* __restrict__ mats, uint64_t numTiles,
                                                  uint64_t loadMap, uint64_t storeMap) {
    extern __shared__ unsigned char smem_raw[];
    cx<double>* lds = reinterpret_cast<cx<double>*>(smem_raw);                 // two remap halves
    double* lmat = reinterpret_cast<double*>(smem_raw + 2 * RT_TILE_BYTES);
    const uint32_t tid = threadIdx.x;         // = lane | wave << 6: the thread's lane and wave slot bits
    uint64_t tile = blockIdx.x;
    if (tile >= numTiles) return;
    // MFMA matrices are copied once, behind the remap area, in the order
    // mat_lds_index reads them (the left half of the real form)
    for (unsigned it = 0; it < numItems; it++) {
        const TileOp& op = items[it];
        if (op.kind != OP_MATK) continue;
        const uint32_t k = op.k, D = 1u << k;
        const double* Mg = reinterpret_cast<const double*>(mats + op.mat_off);
        for (uint32_t e = tid; e < 2 * D * D; e += 512) {
            const uint32_t r = e >> k, c = e & (D - 1);
            lmat[op.lmask + r * D + (c ^ (r & (D - 1)))] = Mg[(size_t)r * 2 * D + c];
        }
    }
    __syncthreads();
    // global address parts: the thread's lane/wave slots and the three register slots
    uint64_t ldLW = 0, stLW = 0;
    for (uint32_t b = 0; b < RT_REG0; b++) {
        if ((tid >> b) & 1u) {
            ldLW |= 1ULL << geom.q[rt_nib(loadMap, b)];
            stLW |= 1ULL << geom.q[rt_nib(storeMap, b)];
        }
    }
    const uint64_t ld0 = 1ULL << geom.q[rt_nib(loadMap, RT_REG0)], ld1 = 1ULL << geom.q[rt_nib(loadMap, RT_REG0 + 1)],
                   ld2 = 1ULL << geom.q[rt_nib(loadMap, RT_REG0 + 2)];
    const uint64_t st0 = 1ULL << geom.q[rt_nib(storeMap, RT_REG0)], st1 = 1ULL << geom.q[rt_nib(storeMap, RT_REG0 + 1)],
                   st2 = 1ULL << geom.q[rt_nib(storeMap, RT_REG0 + 2)];
#define FSA_RT_LD(r) (ldLW | (((r) & 1) ? ld0 : 0ULL) | (((r) & 2) ? ld1 : 0ULL) | (((r) & 4) ? ld2 : 0ULL))
#define FSA_RT_ST(r) (stLW | (((r) & 1) ? st0 : 0ULL) | (((r) & 2) ? st1 : 0ULL) | (((r) & 4) ? st2 : 0ULL))
    cx<double> v[RT_G], nx[PF ? RT_G : 1];
    {
        const uint64_t b0 = ins0_mask(tile, geom.tileMask);
#pragma unroll
        for (int r = 0; r < RT_G; r++) v[r] = a[b0 | FSA_RT_LD(r)];
    }
    uint64_t next = tile + gridDim.x;
    if (PF && next < numTiles) {
        const uint64_t b1 = ins0_mask(next, geom.tileMask);
#pragma unroll
        for (int r = 0; r < RT_G; r++) nx[r] = a[b1 | FSA_RT_LD(r)];
    }
    uint32_t half = 0;                      // remap buffer half in turn (across tiles too)
    for (;;) {
        const uint64_t base = ins0_mask(tile, geom.tileMask);
        for (unsigned it = 0; it < numItems; it++) {
            const TileOp& op = items[it];
            switch (op.kind) {
            case OP_MAT1: {
                if ((base & op.gmask) != op.gmask) break;        // an outside control is 0
                const uint32_t ts = op.p[0];
                if (ts == RT_REG0) rt_mat1_reg<0>(v, op, tid);
                else if (ts == RT_REG0 + 1) rt_mat1_reg<1>(v, op, tid);
                else if (ts == RT_REG0 + 2) rt_mat1_reg<2>(v, op, tid);
                else rt_mat1_lane(v, op, tid, ts);
                break;
            }
            case OP_PAULI: {
                const uint32_t bp = par64(base & op.gmask);
                switch (op.lmask >> RT_REG0) {
                case 0: rt_pauli<0>(v, op, tid, bp); break;
                case 1: rt_pauli<1>(v, op, tid, bp); break;
                case 2: rt_pauli<2>(v, op, tid, bp); break;
                case 3: rt_pauli<3>(v, op, tid, bp); break;
                case 4: rt_pauli<4>(v, op, tid, bp); break;
                case 5: rt_pauli<5>(v, op, tid, bp); break;
                case 6: rt_pauli<6>(v, op, tid, bp); break;
                default: rt_pauli<7>(v, op, tid, bp); break;
                }
                break;
            }
            case OP_PHASE: rt_phase(v, op, tid, par64(base & op.gmask)); break;
            case OP_MATK: {
                const double* M = lmat + op.lmask;
                if (op.k == 2) rt_mfma<2>(v, M);
                else if (op.k == 3) rt_mfma<3>(v, M);
                else if (op.k == 4) rt_mfma<4>(v, M);
                else rt_mfma<5>(v, M);
                break;
            }
            case OP_REMAP:
                rt_remap(v, op, tid, lds + ((size_t)half << 12));
                half ^= 1u;
                break;
            default: break;
            }
        }
        const bool more = next < numTiles;
        if constexpr (PF) {
            // stores of this tile and loads of the tile after the next go out
            // together; the wait makes the next tile's loads land first (vector
            // memory operations complete in order)
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
#pragma unroll
            for (int r = 0; r < RT_G; r++) {
                a[base | FSA_RT_ST(r)] = v[r];
                if (more) v[r] = nx[r];
            }
            if (!more) break;
            const uint64_t after = next + gridDim.x;
            if (after < numTiles) {
                const uint64_t b2 = ins0_mask(after, geom.tileMask);
#pragma unroll
                for (int r = 0; r < RT_G; r++) nx[r] = a[b2 | FSA_RT_LD(r)];
            }
            tile = next;
            next = after;
        } else {
#pragma unroll
            for (int r = 0; r < RT_G; r++) a[base | FSA_RT_ST(r)] = v[r];
            if (!more) break;
            const uint64_t b1 = ins0_mask(next, geom.tileMask);
#pragma unroll
            for (int r = 0; r < RT_G; r++) v[r] = a[b1 | FSA_RT_LD(r)];
            tile = next;
            next += gridDim.x;
        }
    }
#undef FSA_RT_LD
#undef FSA_RT_ST
}

}  // namespace

// FSA_RTILE_PF=0: no register prefetch, two workgroups per CU
bool rtile_prefetch() {
    static const bool pf = [] {
        const char* e = std::getenv("FSA_RTILE_PF");
        return !(e && e[0] == '0');
    }();
    return pf;
}
uint32_t rtile_mat_lds_max() { return RTILE_MAT_LDS_MAX; }

void launch_rtile(hipStream_t s, cx<double>* a, unsigned nq, const TileGeom& geom, const TileOp* ditems,
                  unsigned numItems, const cx<double>* dmats, uint64_t loadMap, uint64_t storeMap, uint32_t matLds) {
    static int numCU = 0;
    if (numCU == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || numCU <= 0)
            numCU = 256;
    }
    if (geom.m != 12) fail(FSA_ERR_INTERNAL, "register-resident passes hold tiles of 2^12 amplitudes");
    const bool pf = rtile_prefetch();
    if (matLds > rtile_mat_lds_max())
        fail(FSA_ERR_INTERNAL, "register-resident pass asks for more LDS for matrices than it has");
    const uint64_t numTiles = 1ULL << (nq - geom.m);
    // prefetching: one workgroup per CU (two tiles in registers, up to 256 VGPRs)
    const uint64_t grid = std::min<uint64_t>(numTiles, (uint64_t)numCU * (pf ? 1u : 2u));
    const size_t lds = 2 * RT_TILE_BYTES + matLds;
    auto* kern = pf ? k_rtile<true> : k_rtile<false>;
    if (lds > 64 * 1024) {
        static std::mutex mu;
        static bool granted[2] = {false, false};
        std::lock_guard<std::mutex> lock(mu);
        if (!granted[pf]) {
            if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(2 * RT_TILE_BYTES + rtile_mat_lds_max())) != hipSuccess)
                fail(FSA_ERR_HIP, "cannot grant the register-resident tile kernel its LDS");
            granted[pf] = true;
        }
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(512), lds, s, a, geom, ditems, numItems, dmats, numTiles,
                       loadMap, storeMap);
}

}  // namespace fsa
