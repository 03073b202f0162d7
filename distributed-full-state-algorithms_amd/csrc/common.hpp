// common.hpp -- types and device index arithmetic shared by the kernels and
// the host side of libfsa.
//
// Layout in HBM: one contiguous array of interleaved complex amplitudes per
// state (complex128 for FP64 states, complex64 for FP32 states), qubit q = bit
// q of the index; density matrices are the column-major 2N-qubit vector.  One
// FP64 amplitude is exactly one 16-byte global_load_dwordx4.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace fsa {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) { cx<R> r; r.x = a; r.y = b; return r; }

template <typename R>
__host__ __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return mk<R>(a.x * s, a.y * s); }

// Amplitude load / store for kernels that stream the state once: FP64 goes
// non-temporal (one 16-byte access; 2-7% faster on gfx950 when each line is
// written whole), FP32 stays cached (non-temporal measured slower there).
template <typename R>
__device__ __forceinline__ cx<R> ld_stream(const cx<R>* p) {
    if constexpr (sizeof(R) == 8) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return mk<R>(v.x, v.y);
    } else {
        return *p;
    }
}
template <typename R>
__device__ __forceinline__ void st_stream(cx<R>* p, cx<R> v) {
    if constexpr (sizeof(R) == 8) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        v2 w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}
// acc + a*b
template <typename R>
__host__ __device__ __forceinline__ cx<R> cfma(cx<R> acc, cx<R> a, cx<R> b) {
    acc.x = acc.x + a.x * b.x - a.y * b.y;
    acc.y = acc.y + a.x * b.y + a.y * b.x;
    return acc;
}

// ---- index arithmetic (reference src/bit_maths.hpp:19-125 as device functions)
// Indices are 64-bit: one MI355X holds up to 2^34 FP64 amplitudes.

// open a zero at position b (insertBit(x, b, 0), bit_maths.hpp:43-49)
__host__ __device__ __forceinline__ uint64_t ins0(uint64_t x, unsigned b) {
    uint64_t low = x & ((1ULL << b) - 1ULL);
    return ((x >> b) << (b + 1)) | low;
}
__host__ __device__ __forceinline__ uint32_t ins0_32(uint32_t x, unsigned b) {
    uint32_t low = x & ((1u << b) - 1u);
    return ((x >> b) << (b + 1)) | low;
}
// open zeros at every set bit of `mask`, ascending (insertBits, bit_maths.hpp:52-59)
__host__ __device__ __forceinline__ uint64_t ins0_mask(uint64_t x, uint64_t mask) {
    while (mask) {
        unsigned b = (unsigned)__builtin_ctzll(mask);
        x = ins0(x, b);
        mask &= mask - 1;
    }
    return x;
}
__host__ __device__ __forceinline__ uint32_t ins0_mask32(uint32_t x, uint32_t mask) {
    while (mask) {
        unsigned b = (unsigned)__builtin_ctz(mask);
        x = ins0_32(x, b);
        mask &= mask - 1;
    }
    return x;
}
// getBitMaskParity (bit_maths.hpp:81-88)
__host__ __device__ __forceinline__ unsigned par64(uint64_t x) { return (unsigned)__builtin_popcountll(x) & 1u; }
__host__ __device__ __forceinline__ unsigned par32(uint32_t x) { return (unsigned)__builtin_popcount(x) & 1u; }

// ---- kernel-argument records -------------------------------------------------

template <typename R>
struct Mat2 { cx<R> m[4]; };                    // row-major 2x2

// Operation kinds executed by the fused tile kernel and the streaming kernels.
enum OpKind : uint32_t {
    OP_MAT1 = 1,     // 2x2 on one target, all-ones control masks
    OP_MATK = 2,     // dense 2^k x 2^k on k targets (matrix in the parameter buffer)
    OP_SWAP = 3,     // swap two qubits
    OP_PAULI = 4,    // (this, other) Pauli tensor/gadget pair map
    OP_PHASE = 5,    // e^{+-i theta} by parity
    OP_DEPHASE1 = 6,
    OP_DEPHASE2 = 7,
    OP_DEPOL1 = 8,
    OP_DEPOL2 = 9,
    OP_DAMP = 10,
    OP_SWEEP = 11,   // tile pass item: a run of ops applied to register groups (see kernels_tile.hip)
    OP_GRAD = 12,    // batched adjoint: accumulate <lambda|P|phi> over the tile (see kernels_batch.hip)
    OP_DIAG4 = 14,   // diagonal factor f[(parity1 << 1) | parity2] (pass-end item applied at the write-back)
    OP_REMAP = 15,   // register-resident passes: move amplitudes to another slot layout (see kernels_rtile.hip)
};

// One item of a fused tile pass, or one micro-op inside a sweep item.
// Item form: positions p[] are bit positions inside the tile (0..m-1); masks
// named l* are tile-local, gmask is a global index mask restricted to qubits
// outside the tile (its parity or all-ones test is fixed for a whole tile).
// OP_SWEEP items: p[0..W-1] = the sweep's tile bits (ascending), k = number of
// micro-ops, mat_off = index of the first micro-op.
// Micro-op form (inside a sweep): p[] and l* masks are sweep-local bit
// indices (0..W-1), gq is a tile-local mask of bits outside the sweep (fixed
// per register group), gmask as above.
struct TileOp {
    uint32_t kind;
    uint32_t k;           // MATK: number of targets; SWEEP: number of micro-ops
    uint32_t p[6];        // positions (targets in caller order for item-form MATK, ascending in micro form)
    uint32_t lmask;       // MAT1: control mask; PAULI: XY mask; PHASE: parity mask;
                          // MATK (MFMA form): offset (reals) of the matrix in the pass's LDS area;
                          // DIAG4: tile-local mask of parity 1 (lmask2: of parity 2; gmask, gq: out-of-tile masks)
    uint32_t lmask2;      // PAULI: YZ mask
    uint64_t gmask;       // MAT1: out-of-tile control mask; PAULI/PHASE: out-of-tile parity mask
    uint64_t gq;          // micro form: the same masks restricted to tile bits outside the sweep
    uint32_t mat_off;     // MATK: matrix offset (complex elements); SWEEP: first micro-op
    uint32_t param;       // adjoint: parameter index of a gadget, or 0xffffffff
    double f[8];          // MAT1: 2x2 matrix; PAULI: this, other*powI; PHASE: e^{i t}, e^{-i t}; channels: constants
};
static_assert(sizeof(TileOp) % 8 == 0, "TileOp must stay 8-byte aligned");

constexpr uint32_t NO_PARAM = 0xffffffffu;
constexpr size_t DRAIN_MAX = 1;            // diagonal items a tile pass applies on its way out of LDS
constexpr uint32_t MAT_UNPLACED = 0xffffffffu;  // MFMA-form MATK item not yet given an LDS offset
// LDS bytes per workgroup for MFMA matrices: what a CU's 160 KiB of LDS leaves
// next to two 64 KiB tiles (16 KiB each) or one 128 KiB tile (32 KiB)
constexpr uint32_t tile_mat_lds_bytes(size_t tileBytes) {
    return tileBytes > 65536 ? (uint32_t)(163840 - tileBytes) : 16384u;
}
constexpr int TILE_MAX_BITS = 14;

// Geometry of one fused pass: the sorted tile qubits (global positions).
struct TileGeom {
    uint8_t q[TILE_MAX_BITS];
    uint8_t m;            // number of tile qubits (tile = 2^m amplitudes)
    uint8_t pad;
    uint64_t tileMask;    // OR of 1 << q[i]
};

// ---- batched circuits (kernels_batch.hip) -------------------------------------
// One job = one circuit on its own n-qubit statevector, started from |0..0>:
// forward plan, lambda = H phi and <H> = <phi|lambda>, then the reverse plan
// (inverse gates on phi and lambda, OP_GRAD items between them).
struct BatchStage {
    uint64_t tileMask;    // global positions of the m tile qubits
    uint32_t m;           // tile bits
    uint32_t item0;       // first item (index into the item array)
    uint32_t nitems;
    uint32_t pad;
};

struct BatchJob {
    uint64_t phi, lam;    // element offsets of phi / lambda in the arena (unused when resident)
    uint32_t n;           // qubits
    uint32_t resident;    // 1: phi and lambda stay in LDS for the whole job (n <= tile bits)
    uint32_t fwd0, nfwd;  // forward stages
    uint32_t rev0, nrev;  // reverse stages
    uint32_t term0, nterm;// observable terms (PauliTermDev)
    uint32_t out;         // result slot of <H> (two doubles)
    uint32_t slot0, nslot;// result slots of the job's gradients, zeroed at the start
    uint32_t nrm;         // result double that receives <phi|phi> after the forward plan
};

}  // namespace fsa
