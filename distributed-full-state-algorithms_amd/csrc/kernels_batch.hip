// kernels_batch.hip -- many independent circuits with their adjoint gradients
// in one persistent kernel.
//
// Each workgroup takes whole jobs (one circuit on one statevector) from a
// global counter, largest first (the host sorts them), and runs the job alone:
//   1. forward plan from |0..0>: fused passes over LDS tiles of phi;
//   2. lambda = H phi and <H> = <phi|lambda> (FP64 sums);
//   3. reverse plan: the inverse gates applied to phi and lambda tile by tile,
//      with OP_GRAD items accumulating <lambda|P_k|phi> in front of the
//      inverse of gadget k (d<H>/d theta_k = -2 Im of it for exp(i theta P)).
// A state of at most M qubits lives in LDS for the whole job (no HBM traffic
// at all); a larger one lives in the arena and every pass streams it through
// LDS one tile at a time.  One workgroup owns a job from start to end, so the
// only synchronisation is the workgroup barrier, every sum has a fixed order
// (results are bit-identical run to run) and there is no inter-workgroup wait:
// each workgroup leaves the loop as soon as the counter passes the last job.
//
// Gadget angles are read from `params` at run time (item field param, sign in
// f[6], i^#Y in f[4..5]), so new angles only need a copy into that array:
// the launch itself can be captured once in a graph and replayed.
#include "launch.hpp"
#include "tile_ops.hpp"

#include <algorithm>
#include <type_traits>

namespace fsa {

namespace {

using namespace tile;

// deposit the low bits of x into the set bits of mask, ascending
__device__ __forceinline__ uint64_t pdep64(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    for (; mask && x; mask &= mask - 1, x >>= 1)
        if (x & 1ULL) r |= mask & (~mask + 1ULL);
    return r;
}

__device__ __forceinline__ double bwave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// workgroup sum of (vr, vi); the total is returned to every thread
template <int TB>
__device__ __forceinline__ void bsum2(double& vr, double& vi, double* red) {
    vr = bwave_sum(vr);
    vi = bwave_sum(vi);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    __syncthreads();                       // red may still be read by a previous sum
    if (lane == 0) { red[2 * wave] = vr; red[2 * wave + 1] = vi; }
    __syncthreads();
    double a = 0, b = 0;
#pragma unroll
    for (int w = 0; w < TB / 64; w++) { a += red[2 * w]; b += red[2 * w + 1]; }
    vr = a;
    vi = b;
}

// gadget factors from the run-time angle (forward: sign +1, inverse: -1)
__device__ __forceinline__ void bind_param(TileOp& op, const double* __restrict__ params) {
    if (op.param == NO_PARAM) return;
    const double t = op.f[6] * params[op.param];
    double s, c;
    sincos(t, &s, &c);
    if (op.kind == OP_PAULI) {            // exp(i t P): this = cos t, other = i sin t * i^#Y
        op.f[0] = c; op.f[1] = 0.0;
        op.f[2] = -s * op.f[5]; op.f[3] = s * op.f[4];
    } else {                               // OP_PHASE: e^{+i t} even parity, e^{-i t} odd
        op.f[0] = c; op.f[1] = s;
        op.f[2] = c; op.f[3] = -s;
    }
}

// <lambda|P|phi> over one tile: sum_l conj(lam_l) i^#Y (-1)^{par(k & myz)} phi_k, k = l ^ mxy
template <typename R, int TB>
__device__ __forceinline__ void grad_partial(const cx<R>* phi, const cx<R>* lam, uint32_t T, uint64_t base,
                                             const TileOp& op, double& vr, double& vi) {
    const uint32_t bp = par64(base & op.gmask);
    double ar = 0, ai = 0;
    for (uint32_t l = threadIdx.x; l < T; l += TB) {
        const uint32_t k = l ^ op.lmask;
        const cx<R> x = lam[l], y = phi[k];
        const double sg = (bp ^ par32(k & op.lmask2)) ? -1.0 : 1.0;
        ar += sg * ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
        ai += sg * ((double)x.x * (double)y.y - (double)x.y * (double)y.x);
    }
    vr = op.f[0] * ar - op.f[1] * ai;
    vi = op.f[0] * ai + op.f[1] * ar;
}

// A gradient item fused with the inverse gadget after it (same Pauli masks):
// one sweep over the pairs (l, l ^ mxy) of both tiles accumulates the item's
// <lambda|P|phi> terms (as grad_partial) and applies the gadget `o` to phi and
// lambda with the arithmetic of tile_apply's OP_PAULI / OP_PHASE.
template <typename R, int TB>
__device__ __forceinline__ void grad_apply(cx<R>* phi, cx<R>* lam, uint32_t T, uint64_t base, const TileOp& g,
                                           const TileOp& o, double& vr, double& vi) {
    const uint32_t bp = par64(base & g.gmask);
    const uint32_t lxy = g.lmask, lyz = g.lmask2;
    double ar = 0, ai = 0;
    auto term = [&](const cx<R>& x, const cx<R>& y, uint32_t k) {   // conj(x) y (-1)^{par(k & myz)}
        const double sg = (bp ^ par32(k & lyz)) ? -1.0 : 1.0;
        ar += sg * ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
        ai += sg * ((double)x.x * (double)y.y - (double)x.y * (double)y.x);
    };
    if (o.kind == OP_PHASE) {
        const cx<double> f0 = ldc<double>(o.f, 0), f1 = ldc<double>(o.f, 1);
        const uint32_t bq = par64(base & o.gmask);
        for (uint32_t l = threadIdx.x; l < T; l += TB) {
            const cx<R> x = lam[l], y = phi[l];
            term(x, y, l);
            const cx<double> f = (bq ^ par32(l & o.lmask)) ? f1 : f0;
            phi[l] = dn<R>(cmul(up<double>(y), f));
            lam[l] = dn<R>(cmul(up<double>(x), f));
        }
    } else {
        const cx<double> th = ldc<double>(o.f, 0), fe = ldc<double>(o.f, 1);
        const cx<double> fo = mk<double>(-fe.x, -fe.y);
        const uint32_t bq = par64(base & o.gmask);
        if (lxy == 0) {
            for (uint32_t l = threadIdx.x; l < T; l += TB) {
                const cx<R> x = lam[l], y = phi[l];
                term(x, y, l);
                const cx<double> f = (bq ^ par32(l & o.lmask2)) ? fo : fe;
                const cx<double> yd = up<double>(y), xd = up<double>(x);
                phi[l] = dn<R>(cfma(cmul(th, yd), f, yd));
                lam[l] = dn<R>(cfma(cmul(th, xd), f, xd));
            }
        } else {
            const unsigned hb = 31u - (unsigned)__builtin_clz(lxy);
            for (uint32_t j = threadIdx.x; j < (T >> 1); j += TB) {
                const uint32_t i0 = ins0_32(j, hb), i1 = i0 ^ lxy;
                const cx<R> p0 = phi[i0], p1 = phi[i1], l0 = lam[i0], l1 = lam[i1];
                term(l0, p1, i1);
                term(l1, p0, i0);
                const cx<double> f0 = (bq ^ par32(i0 & o.lmask2)) ? fo : fe;
                const cx<double> f1 = (bq ^ par32(i1 & o.lmask2)) ? fo : fe;
                const cx<double> q0 = up<double>(p0), q1 = up<double>(p1), m0 = up<double>(l0), m1 = up<double>(l1);
                phi[i0] = dn<R>(cfma(cmul(th, q0), f1, q1));
                phi[i1] = dn<R>(cfma(cmul(th, q1), f0, q0));
                lam[i0] = dn<R>(cfma(cmul(th, m0), f1, m1));
                lam[i1] = dn<R>(cfma(cmul(th, m1), f0, m0));
            }
        }
    }
    vr = g.f[0] * ar - g.f[1] * ai;
    vi = g.f[0] * ai + g.f[1] * ar;
}

// lambda_j = sum_t f_t (-1)^{par(k & myz_t)} phi_k, k = j ^ mxy_t, for j in [j0, j0 + T);
// accumulates <phi|lambda> over those j
template <typename R, int TB>
__device__ __forceinline__ void apply_terms(const cx<R>* phiAll, cx<R>* lamOut, uint64_t j0, uint32_t T,
                                            const PauliTermDev* __restrict__ terms, uint32_t nterm,
                                            double& er, double& ei, double& nn) {
    for (uint32_t l = threadIdx.x; l < T; l += TB) {
        const uint64_t j = j0 + l;
        double ar = 0, ai = 0;
        for (uint32_t t = 0; t < nterm; t++) {
            const PauliTermDev term = terms[t];
            const uint64_t k = j ^ term.mxy;
            const double sg = par64(k & term.myz) ? -1.0 : 1.0;
            const cx<R> v = phiAll[k];
            const double fr = sg * term.fr, fi = sg * term.fi;
            ar += fr * (double)v.x - fi * (double)v.y;
            ai += fr * (double)v.y + fi * (double)v.x;
        }
        const cx<R> p = phiAll[j];
        nn += (double)p.x * (double)p.x + (double)p.y * (double)p.y;
        er += (double)p.x * ar + (double)p.y * ai;
        ei += (double)p.x * ai - (double)p.y * ar;
        lamOut[l] = mk<R>((R)ar, (R)ai);
    }
}

// Gradient items of one stage accumulate in LDS (gacc, indexed by the item's
// p[0] = its index among the stage's gradient items) over the tiles a
// workgroup processes, and are written out once per stage (flush_grads).
constexpr uint32_t MAX_STAGE_GRADS = 512;

// the items of one stage on the LDS tile(s); OP_GRAD items add their tile sum
// into gacc (tiles in a fixed order)
template <typename R, int TB, bool REV, int ROWS>
__device__ __forceinline__ void run_items(cx<R>* phi, cx<R>* lam, uint32_t T, uint64_t base,
                                          const TileOp* __restrict__ items, uint32_t i0, uint32_t ni,
                                          const cx<double>* __restrict__ mats, const double* __restrict__ params,
                                          double* gacc, double* red, bool firstTile) {
    for (uint32_t i = i0; i < i0 + ni; i++) {
        TileOp op = items[i];
        if (REV && op.kind == OP_GRAD) {
            double vr, vi;
            const bool fused = op.k == 1u;  // the next item is this gadget's inverse
            if (fused) {
                TileOp inv = items[i + 1];
                bind_param(inv, params);
                grad_apply<R, TB>(phi, lam, T, base, op, inv, vr, vi);
                i++;
            } else {
                grad_partial<R, TB>(phi, lam, T, base, op, vr, vi);
            }
            bsum2<TB>(vr, vi, red);
            if (threadIdx.x == 0) {
                double* r = gacc + 2 * op.p[0];
                if (firstTile) { r[0] = vr; r[1] = vi; }
                else { r[0] += vr; r[1] += vi; }
            }
            if (fused) __syncthreads();     // the tiles were written (a gradient sum alone only reads them)
            continue;
        }
        bind_param(op, params);
        // full tiles (T = ROWS * TB) load batches of four pairs before updating them
        // (one workgroup per CU: registers to spare)
        tile_apply<R, TB, 1, 1, double, ROWS, 4>(phi, T, base, op, mats);
        if (REV) tile_apply<R, TB, 1, 1, double, ROWS, 4>(lam, T, base, op, mats);
        __syncthreads();
    }
}

// gradient sums of a finished stage -> res slots (item field mat_off)
template <int TB>
__device__ __forceinline__ void flush_grads(const BatchStage& st, const TileOp* __restrict__ items,
                                            const double* gacc, double* __restrict__ res) {
    __syncthreads();
    for (uint32_t i = st.item0 + threadIdx.x; i < st.item0 + st.nitems; i += TB) {
        if (items[i].kind == OP_GRAD) {
            const uint32_t g = items[i].p[0];
            const size_t slot = items[i].mat_off;
            res[2 * slot] = gacc[2 * g];
            res[2 * slot + 1] = gacc[2 * g + 1];
        }
    }
    __syncthreads();
}

template <typename R, int TB, int M>
__global__ __launch_bounds__(TB) void k_batch(const BatchJob* __restrict__ jobs, unsigned numJobs,
                                              const BatchStage* __restrict__ stages,
                                              const TileOp* __restrict__ items, const cx<double>* __restrict__ mats,
                                              const PauliTermDev* __restrict__ terms,
                                              const double* __restrict__ params, cx<R>* __restrict__ arena,
                                              double* __restrict__ res, unsigned* __restrict__ counter) {
    constexpr uint32_t TMAX = 1u << M;
    constexpr uint32_t KMAX = TMAX / TB > 0 ? TMAX / TB : 1;
    __shared__ cx<R> lds[2 * TMAX];
    __shared__ uint64_t offHi[KMAX];
    __shared__ double red[2 * (TB / 64)];
    __shared__ double gacc[2 * MAX_STAGE_GRADS];
    __shared__ unsigned jobIdx;
    const uint32_t tid = threadIdx.x;

    for (;;) {
        __syncthreads();
        if (tid == 0) jobIdx = atomicAdd(counter, 1u);
        __syncthreads();
        const unsigned jx = jobIdx;
        if (jx >= numJobs) break;
        const BatchJob J = jobs[jx];

        if (J.resident) {
            // ---- whole job in LDS: phi = lds[0, T), lambda = lds[T, 2T)
            const uint32_t T = 1u << J.n;
            cx<R>* phi = lds;
            cx<R>* lam = lds + T;
            for (uint32_t l = tid; l < T; l += TB) phi[l] = mk<R>(l == 0 ? (R)1 : (R)0, (R)0);
            __syncthreads();
            for (uint32_t s = J.fwd0; s < J.fwd0 + J.nfwd; s++)
                run_items<R, TB, false, KMAX>(phi, lam, T, 0, items, stages[s].item0, stages[s].nitems, mats, params, gacc,
                                        red, true);
            double er = 0, ei = 0, nn = 0, nz = 0;
            apply_terms<R, TB>(phi, lam, 0, T, terms + J.term0, J.nterm, er, ei, nn);
            bsum2<TB>(er, ei, red);
            bsum2<TB>(nn, nz, red);
            if (tid == 0) {
                res[2 * (size_t)J.out] = er;
                res[2 * (size_t)J.out + 1] = ei;
                res[J.nrm] = nn;
            }
            __syncthreads();
            for (uint32_t s = J.rev0; s < J.rev0 + J.nrev; s++) {
                const BatchStage st = stages[s];
                run_items<R, TB, true, KMAX>(phi, lam, T, 0, items, st.item0, st.nitems, mats, params, gacc, red, true);
                flush_grads<TB>(st, items, gacc, res);
            }
            continue;
        }

        // ---- state in the arena, streamed through LDS tile by tile
        cx<R>* gphi = arena + J.phi;
        cx<R>* glam = arena + J.lam;
        const uint64_t all = (J.n >= 64) ? ~0ULL : ((1ULL << J.n) - 1ULL);
        // one streamed pass of stage s; REV: the reverse sweep (phi and lambda)
        auto stream_pass = [&](auto revTag, uint32_t s, bool fromZero) {
            constexpr bool REV = decltype(revTag)::value;
            const BatchStage st = stages[s];
            const uint32_t T = 1u << st.m;
            const uint32_t K = T > TB ? T / TB : 1u;
            const uint64_t outMask = all & ~st.tileMask;
            const uint64_t numTiles = 1ULL << (J.n - st.m);
            __syncthreads();
            // every row gets an offset; rows past K alias in-tile addresses (pdep
            // drops the bits above the tile), so the prefetch needs no row guard
            if (tid < KMAX) offHi[tid] = pdep64((uint64_t)tid * TB, st.tileMask);
            const uint64_t offLo = pdep64(tid, st.tileMask);
            __syncthreads();
            // Tiles of one pass are disjoint: while the items of tile t run in
            // LDS, the loads of tile t + 1 (registers) and the write-back of tile
            // t - 1 are in flight.  The loads are unconditional (a value merged
            // from branches is copied, and so waited for, right after its load);
            // the first forward pass, which starts from |0..0>, reads the arena
            // and ignores what it read.  A tile here has T >= 2^10 >= TB
            // amplitudes, K = T / TB <= KMAX per thread.
            R px[KMAX], py[KMAX], qx[KMAX], qy[KMAX];
            auto load = [&](uint64_t tt) {
                const uint64_t b = pdep64(tt, outMask);
#pragma unroll
                for (uint32_t k = 0; k < KMAX; k++) {
                    const cx<R> v = gphi[b | offLo | offHi[k]];
                    px[k] = v.x;
                    py[k] = v.y;
                    if constexpr (REV) {
                        const cx<R> w = glam[b | offLo | offHi[k]];
                        qx[k] = w.x;
                        qy[k] = w.y;
                    }
                }
            };
            // amplitude of row k at global index g: |0..0> on the first pass, else loaded
            auto phiAt = [&](uint32_t k, uint64_t g) {
                return fromZero ? mk<R>(g == 0 ? (R)1 : (R)0, (R)0) : mk<R>(px[k], py[k]);
            };
            load(0);
#pragma unroll
            for (uint32_t k = 0; k < KMAX; k++) {
                if (k < K) {
                    lds[tid + k * TB] = phiAt(k, offLo | offHi[k]);
                    if constexpr (REV) lds[TMAX + tid + k * TB] = mk<R>(qx[k], qy[k]);
                }
            }
            __syncthreads();
            if (1 < numTiles) load(1);
            for (uint64_t t = 0; t < numTiles; t++) {
                const uint64_t base = pdep64(t, outMask);
                run_items<R, TB, REV, KMAX>(lds, lds + TMAX, T, base, items, st.item0, st.nitems, mats, params, gacc,
                                      red, t == 0);
                __syncthreads();   // the last item may be a gradient sum still reading other rows
                // Each thread writes its own rows of tile t back and fills them
                // with tile t + 1; the wait makes tile t + 1's loads land before the
                // first store of tile t is issued (vector memory operations
                // complete in order).
                __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
                const bool more = t + 1 < numTiles;
                const uint64_t nbase = more ? pdep64(t + 1, outMask) : 0;
#pragma unroll
                for (uint32_t k = 0; k < KMAX; k++) {
                    if (k < K) {
                        const uint32_t l = tid + k * TB;
                        const uint64_t o = offLo | offHi[k];
                        const R vx = lds[l].x, vy = lds[l].y;
                        if (more) lds[l] = phiAt(k, nbase | o);
                        gphi[base | o] = mk<R>(vx, vy);
                        if constexpr (REV) {
                            const R wx = lds[TMAX + l].x, wy = lds[TMAX + l].y;
                            if (more) lds[TMAX + l] = mk<R>(qx[k], qy[k]);
                            glam[base | o] = mk<R>(wx, wy);
                        }
                    }
                }
                if (!more) break;
                if (t + 2 < numTiles) load(t + 2);
                __syncthreads();   // tile t + 1 is in LDS
            }
            __syncthreads();   // the pass is in the arena before anything reads it
            if constexpr (REV) flush_grads<TB>(st, items, gacc, res);
        };
        for (int dir = 0; dir < 2; dir++) {
            const uint32_t s0 = dir == 0 ? J.fwd0 : J.rev0;
            const uint32_t ns = dir == 0 ? J.nfwd : J.nrev;
            for (uint32_t s = s0; s < s0 + ns; s++) {
                if (dir == 0) stream_pass(std::false_type{}, s, s == s0);
                else stream_pass(std::true_type{}, s, false);
            }
            if (dir == 0) {
                // the forward pass is complete in the arena (this workgroup's own
                // stores, ordered by the barrier above): lambda = H phi, <H>
                double er = 0, ei = 0, nn = 0, nz = 0;
                const uint64_t num = 1ULL << J.n;
                for (uint64_t j0 = 0; j0 < num; j0 += TMAX) {
                    const uint32_t T = (uint32_t)((num - j0) < TMAX ? (num - j0) : TMAX);
                    apply_terms<R, TB>(gphi, glam + j0, j0, T, terms + J.term0, J.nterm, er, ei, nn);
                }
                bsum2<TB>(er, ei, red);
                bsum2<TB>(nn, nz, red);
                if (tid == 0) {
                    res[2 * (size_t)J.out] = er;
                    res[2 * (size_t)J.out + 1] = ei;
                    res[J.nrm] = nn;
                }
                __syncthreads();
            }
        }
    }
}

__global__ void k_batch_reset(unsigned* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

}  // namespace

template <typename R, int TB, int M>
static const void* batch_kernel_ptr() { return reinterpret_cast<const void*>(&k_batch<R, TB, M>); }

template <typename R>
int batch_max_bits() { return sizeof(R) == 8 ? 12 : 13; }

// Workgroups of one configuration that fit on a CU at once.
template <typename R>
int batch_occupancy(int tb, int m) {
    int blocks = 0;
    const void* f = nullptr;
#define FSA_BK(TBV, MV) if (tb == TBV && m == MV) f = batch_kernel_ptr<R, TBV, MV>();
    FSA_BK(256, 10) FSA_BK(256, 11) FSA_BK(256, 12) FSA_BK(512, 10) FSA_BK(512, 11) FSA_BK(512, 12)
    if constexpr (sizeof(R) == 4) { FSA_BK(256, 13) FSA_BK(512, 13) }
#undef FSA_BK
    if (!f) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, tb, 0) != hipSuccess) return 0;
    return blocks;
}

template <typename R>
bool launch_batch(hipStream_t s, int tb, int m, unsigned grid, const BatchJob* jobs, unsigned numJobs,
                  const BatchStage* stages, const TileOp* items, const cx<double>* mats, const PauliTermDev* terms,
                  const double* params, cx<R>* arena, double* res, unsigned* counter) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, counter);
#define FSA_BL(TBV, MV)                                                                                    \
    if (tb == TBV && m == MV) {                                                                            \
        hipLaunchKernelGGL((k_batch<R, TBV, MV>), dim3(grid), dim3(TBV), 0, s, jobs, numJobs, stages, items, \
                           mats, terms, params, arena, res, counter);                                      \
        return true;                                                                                       \
    }
    FSA_BL(256, 10) FSA_BL(256, 11) FSA_BL(256, 12) FSA_BL(512, 10) FSA_BL(512, 11) FSA_BL(512, 12)
    if constexpr (sizeof(R) == 4) { FSA_BL(256, 13) FSA_BL(512, 13) }
#undef FSA_BL
    return false;
}

template int batch_max_bits<double>();
template int batch_max_bits<float>();
template int batch_occupancy<double>(int, int);
template int batch_occupancy<float>(int, int);
template bool launch_batch<double>(hipStream_t, int, int, unsigned, const BatchJob*, unsigned, const BatchStage*,
                                   const TileOp*, const cx<double>*, const PauliTermDev*, const double*,
                                   cx<double>*, double*, unsigned*);
template bool launch_batch<float>(hipStream_t, int, int, unsigned, const BatchJob*, unsigned, const BatchStage*,
                                  const TileOp*, const cx<double>*, const PauliTermDev*, const double*, cx<float>*,
                                  double*, unsigned*);

}  // namespace fsa
