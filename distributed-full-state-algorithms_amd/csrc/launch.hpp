// launch.hpp -- host-side launchers of the HIP kernels (internal to libfsa).
#pragma once

#include "common.hpp"

#include <algorithm>
#include <complex>

namespace fsa {

constexpr int BLOCK = 256;

// The block's share of the work (streaming kernels, and the grid-stride
// loops of the FP32 adjoint sweep kernels).  Blocks are dispatched round-robin
// over the 8 XCDs (block b to XCD b % 8); this order gives each XCD a contiguous eighth
// of the shares instead of every eighth one.  At 30 qubits FP64 it streams
// 3-8% faster for most targets (phase gadget 5.69 -> 6.05 TB/s, swap 5.1-5.6
// -> 5.9-6.0, one-qubit gate on qubits 3-10 5.4 -> 5.8) and 1-4% slower for a
// few (one-qubit gate on qubit 18, 20, 24 or 28).  Every block index still
// maps to one share, so the kernels' bounds checks are unchanged; per-block
// partial sums keep the physical block's slot.
__device__ __forceinline__ uint64_t stream_block() {
    const uint32_t G = gridDim.x;
    if (G % 8u) return blockIdx.x;
    return (uint64_t)(blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
}

template <typename R> struct Offs32 { uint64_t o[32]; };

// ---- streaming kernels: one HBM pass per operation ---------------------------
template <typename R>
void launch_mat1(hipStream_t s, cx<R>* a, unsigned nq, unsigned t, uint64_t ctrlMask, const Mat2<R>& g);
template <typename R>
void launch_swap(hipStream_t s, cx<R>* a, unsigned nq, unsigned q1, unsigned q2);
// dense k-target gate, matrix (device, row-major) dmat; k <= 12
template <typename R>
void launch_matk(hipStream_t s, cx<R>* a, unsigned nq, const unsigned* targets, unsigned k,
                 const cx<R>* dmat, const cx<R>* dmatT);

// dense k-target gate, 6 <= k <= 12, as an MFMA GEMM over column blocks of
// groups (kernels_gemm.hip).  The matrix is packed in MFMA fragment order
// (gemm_pack_matrix); gates whose row panels span several workgroups need a
// scratch area of gemm_scratch_elems amplitudes.
template <typename R> struct alignas(2 * sizeof(R)) GemmFrag { R re, im; };
constexpr uint32_t GEMM_CHUNK = 4096;   // amplitudes a workgroup stages at a time
struct GemmGeom {
    uint64_t fixmask;          // target bits and column bits (the column block index skips them)
    uint32_t k, kcBits, cbits; // targets; log2 rows per staged chunk; log2 columns per block
    uint32_t rowUnits, rbw;    // waves along the rows; 16-row blocks per wave
    uint32_t panels;           // workgroups per column block (row panels)
    uint32_t inplace;          // one panel and one chunk: results go straight back
    uint8_t tq[12];            // targets in matrix order (bit i of a row index is qubit tq[i])
    uint8_t lpos[16];          // chunk tile bits, ascending: qubit of local bit b ...
    uint32_t lw[16];           // ... and its weight in the LDS slot (row * CBW + column)
    uint8_t fpos[20];          // full tile bits (every target and the column bits), ascending
    uint32_t fw[20];
};
bool gemm_fits(unsigned nq, unsigned k);
GemmGeom gemm_geom(unsigned nq, const unsigned* targets, unsigned k, size_t realBytes);
uint64_t gemm_scratch_elems(const GemmGeom& g, unsigned nq, size_t realBytes);
template <typename R>
void gemm_pack_matrix(const std::complex<double>* m, unsigned k, GemmFrag<R>* out);
template <typename R>
void launch_matk_gemm(hipStream_t s, cx<R>* a, unsigned nq, const GemmGeom& g, const GemmFrag<R>* dfrag,
                      cx<R>* scratch);

template <typename R>
void launch_pauli(hipStream_t s, cx<R>* a, unsigned nq, uint64_t mxy, uint64_t myz, cx<R> thisFac,
                  cx<R> evenFac);
template <typename R>
void launch_phase(hipStream_t s, cx<R>* a, unsigned nq, uint64_t mask, cx<R> f0, cx<R> f1);
template <typename R>
void launch_dephase1(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R fac);
template <typename R>
void launch_dephase2(hipStream_t s, cx<R>* a, unsigned N, unsigned q1, unsigned q2, R term);
template <typename R>
void launch_depol1(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R c1, R c2, R c3);
template <typename R>
void launch_depol2(hipStream_t s, cx<R>* a, unsigned nq, const unsigned* q /* [4] */, R c1, R c2, R c3);
template <typename R>
void launch_damp(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R prob, R c1, R c2);
template <typename R>
void launch_scale(hipStream_t s, cx<R>* a, uint64_t num, R f);

// ---- fused tile pass ----------------------------------------------------------
// items: plain ops or OP_SWEEP items whose micro-ops live in dmicro; flags
// bits 2-3: kernel variant (0 simple ops, 1 adds VALU dense gates and
// two-qubit depolarising, 2 MFMA dense gates with matrices in LDS, 3 with
// matrices in HBM), bit 4: the pass has sweeps, bits 8-15: the number of
// trailing OP_DIAG4 items (0 or 1) applied in registers as each tile is
// written back.
// matLds: bytes of LDS after the tile for MFMA matrices (variants 2 and 3).
template <typename R>
void launch_tile(hipStream_t s, cx<R>* a, unsigned nq, const TileGeom& geom, const TileOp* ditems,
                 unsigned numItems, const TileOp* dmicro, const cx<R>* dmats, uint32_t flags, uint32_t matLds = 0);
// register-resident pass over FP64 tiles of 2^12 amplitudes (kernels_rtile.hip):
// items in slot space, loadMap / storeMap the slot layouts of the load and
// the store, matLds the LDS bytes of the pass's MFMA matrices
void launch_rtile(hipStream_t s, cx<double>* a, unsigned nq, const TileGeom& geom, const TileOp* ditems,
                  unsigned numItems, const cx<double>* dmats, uint64_t loadMap, uint64_t storeMap, uint32_t matLds);
constexpr uint32_t RTILE_MAT_LDS_MAX = 163840 - 2 * 65536;   // one workgroup per CU: LDS beyond the two remap halves
bool rtile_prefetch();              // the pass kernel keeps the next tile in registers (one workgroup per CU)
uint32_t rtile_mat_lds_max();       // LDS bytes a register-resident pass has for MFMA matrices
// largest tile (log2 amplitudes) the LDS budget allows for this precision
template <typename R> constexpr int tile_bits_max() { return sizeof(R) == 8 ? 12 : 13; }
// largest tile a pass may use when asked for (FSA_TILE_BITS): 128 KiB of LDS
template <typename R> constexpr int tile_bits_limit() { return sizeof(R) == 8 ? 13 : 13; }
// register-sweep width: 2^W amplitudes per thread = full tile / 512 threads
template <typename R> constexpr int sweep_bits() { return sizeof(R) == 8 ? 2 : 4; }

// ---- init / copy / conversion ------------------------------------------------
template <typename R>
void launch_init_gaussian(hipStream_t s, cx<R>* a, unsigned nq, uint64_t seed);
template <typename R>
void launch_init_basis(hipStream_t s, cx<R>* a, uint64_t num, uint64_t index);
template <typename R>
void launch_scale_by_inv_sqrt(hipStream_t s, cx<R>* a, uint64_t num, const double* dnorm2);
void launch_f64_to_f32(hipStream_t s, const cx<double>* in, cx<float>* out, uint64_t num);
void launch_f32_to_f64(hipStream_t s, const cx<float>* in, cx<double>* out, uint64_t num);

// ---- reductions (FP64 accumulation, fixed order: deterministic) --------------
// partial-sum workspace requirement in doubles for a reduction over `num` items
uint64_t reduce_workspace_doubles(uint64_t num);
template <typename R>
void launch_norm2(hipStream_t s, const cx<R>* a, uint64_t num, double* ws, double* dout);
template <typename R>
void launch_inner(hipStream_t s, const cx<R>* a, const cx<R>* b, uint64_t num, double* ws, double* dout);
// sum over terms of coeff_t * (-i)^{nY_t} * sum_r sign * rho[r + 2^N (r ^ mxy_t)]
struct PauliTermDev { uint64_t mxy, myz; double fr, fi; };   // factor f = coeff * phase
uint64_t expec_workspace_doubles(unsigned nbits, unsigned numTerms);
// one adjoint reverse step for a parametrised gadget (kernels_reduce.hip)
struct AdjointStepDev {
    uint64_t mxy, myz;            // Pauli masks (phase gadget: mxy = 0, myz = its parity mask)
    double fr, fi;                // factor of the term <lambda| f P |phi> (i^{#Y}; 1 for phase gadgets)
    double thr, thi, fer, fei;    // Pauli gadget U^dag: v_j <- th v_j + (+-fe) v_{j ^ mxy}
    double e0r, e0i, e1r, e1i;    // phase gadget U^dag: v_j <- v_j e0 (even parity) or e1 (odd)
    uint32_t phase;               // 1: phase gadget
    uint32_t slot;                // row slot of its partial sums (the gadget's parameter)
};
// Partial sums of the adjoint sweep: every launch below writes, for each of
// its gadgets, one (re, im) row per workgroup at rows + stride * slot + 2 *
// workgroup, and returns the number of workgroups; launch_grad_sums adds each
// slot's rows in row order once the sweep is done (one launch for all of them).
uint64_t adjoint_row_stride(unsigned nq);   // doubles between slots
template <typename R>
unsigned launch_adjoint_step(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointStepDev& st,
                             double* rows, uint64_t stride);
// res[2 p], res[2 p + 1] = sum of the first nb[p] rows of slot p, p < numSlots
void launch_grad_sums(hipStream_t s, const double* rows, uint64_t stride, const uint32_t* nb, unsigned numSlots,
                      double* res);
// a run of consecutive diagonal gadgets of the reverse sweep in one pass
// (kernels_reduce.hip): every term <lambda|f_g P_g|phi> is taken from the
// run's input states (diagonal unitaries leave <D lambda|P|D phi> unchanged
// for diagonal P), then every gadget's U^dag factor is applied to both states
constexpr unsigned ADJ_DIAG_MAX = 8;
struct AdjointDiagDev {
    uint64_t myz[ADJ_DIAG_MAX];                       // parity masks
    double fr[ADJ_DIAG_MAX], fi[ADJ_DIAG_MAX];        // term factors f_g
    double e0r[ADJ_DIAG_MAX], e0i[ADJ_DIAG_MAX];      // U^dag factor on even parity
    double e1r[ADJ_DIAG_MAX], e1i[ADJ_DIAG_MAX];      // and on odd parity
    uint32_t slot[ADJ_DIAG_MAX];                      // row slots of the gadgets' partial sums
    uint32_t count;                                   // gadgets in the run (1..ADJ_DIAG_MAX)
    uint32_t pad;
};
// a run of 2-3 one-qubit X/Y gadgets on distinct qubits in one pass: each
// commutes with the others' Pauli operators, so every term is taken from the
// run's input states; then each gadget's U^dag acts on its pairs
constexpr unsigned ADJ_QRUN_MAX = 3;
struct AdjointQubitRunDev {
    uint32_t q[ADJ_QRUN_MAX];                         // qubit of gadget g (run order)
    uint32_t y[ADJ_QRUN_MAX];                         // 1: Y on it (signs by its value), 0: X
    uint32_t sorted[ADJ_QRUN_MAX];                    // the run's qubits ascending
    uint32_t slot[ADJ_QRUN_MAX];                      // row slots of the gadgets' partial sums
    uint32_t count;                                   // gadgets in the run (2..ADJ_QRUN_MAX)
    double fr[ADJ_QRUN_MAX], fi[ADJ_QRUN_MAX];        // term factors f_g = i^{#Y}
    double thr[ADJ_QRUN_MAX], thi[ADJ_QRUN_MAX];      // U^dag: v_j <- th v_j + (+-fe) v_{j ^ bit}
    double fer[ADJ_QRUN_MAX], fei[ADJ_QRUN_MAX];
};
// tail: a diagonal run that follows in the sweep, applied in the same pass
// (only for runs whose qubits are all below 6; nullptr for none)
template <typename R>
unsigned launch_adjoint_qrun(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointQubitRunDev& st,
                             double* rows, uint64_t stride, const AdjointDiagDev* tail = nullptr);
template <typename R>
unsigned launch_adjoint_diag(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointDiagDev& st,
                             double* rows, uint64_t stride);
template <typename R>
void launch_pauli_element(hipStream_t s, const cx<R>* a, const cx<R>* b, uint64_t num, PauliTermDev term,
                          double* ws, double* dout);
template <typename R>
void launch_dm_expec(hipStream_t s, const cx<R>* rho, unsigned N, const PauliTermDev* dterms,
                     unsigned numTerms, double* ws, double* dout);
template <typename R>
void launch_sv_expec(hipStream_t s, const cx<R>* psi, unsigned n, const PauliTermDev* dterms,
                     unsigned numTerms, double* ws, double* dout);
template <typename R>
void launch_apply_pauli_sum(hipStream_t s, const cx<R>* in, cx<R>* out, unsigned n,
                            const PauliTermDev* dterms, unsigned numTerms);
template <typename R>
void launch_partial_trace(hipStream_t s, const cx<R>* in, cx<R>* out, unsigned N,
                          const unsigned* sortedTargets, unsigned k);

}  // namespace fsa
