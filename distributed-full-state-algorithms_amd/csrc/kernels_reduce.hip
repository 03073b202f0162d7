// kernels_reduce.hip -- reductions (FP64 accumulation, fixed order), partial
// trace, Pauli-sum application, state initialisation and precision conversion.
//
// Reductions are two-level and deterministic: a fixed grid writes one partial
// per workgroup (wavefront shuffle, then LDS across the 4 waves), and a single
// workgroup sums the partials in index order.  Results never depend on timing.
#include "engine.hpp"

#include <cstring>
#include "launch.hpp"

namespace fsa {

namespace {

constexpr int RED_ITEMS = 8;
// grid cap of the grid-stride reductions and adjoint sweep kernels: four
// 256-thread blocks per CU.  The 24-qubit adjoint ran 1.3-1.4% faster than with
// 2048 (FP64 diagonal passes 4% faster) and 4096 was slower (same box).
constexpr unsigned RED_MAX_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum of (vr, vi) over the workgroup into lane 0 of wave 0
__device__ __forceinline__ void block_sum2(double& vr, double& vi) {
    __shared__ double sr[BLOCK / 64], si[BLOCK / 64];
    vr = wave_sum(vr);
    vi = wave_sum(vi);
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) { sr[wave] = vr; si[wave] = vi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < BLOCK / 64; w++) { a += sr[w]; b += si[w]; }
        vr = a; vi = b;
    }
}

unsigned red_blocks(uint64_t num) {
    uint64_t per = (uint64_t)BLOCK * RED_ITEMS;
    uint64_t b = (num + per - 1) / per;
    if (b < 1) b = 1;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    return (unsigned)b;
}

__global__ __launch_bounds__(BLOCK) void k_final_sum(const double* __restrict__ part, unsigned numParts,
                                                     double* __restrict__ out) {
    double vr = 0, vi = 0;
    for (unsigned i = threadIdx.x; i < numParts; i += BLOCK) { vr += part[2 * i]; vi += part[2 * i + 1]; }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) { out[0] = vr; out[1] = vi; }
}

// per-slot sums of the adjoint sweep's partial rows: block p adds the first
// nb[p] rows of slot p in row order (the order k_final_sum uses)
__global__ __launch_bounds__(BLOCK) void k_grad_sums(const double* __restrict__ rows, uint64_t stride,
                                                     const uint32_t* __restrict__ nb, double* __restrict__ res) {
    const double* r = rows + stride * blockIdx.x;
    const unsigned n = nb[blockIdx.x];
    double vr = 0, vi = 0;
    for (unsigned i = threadIdx.x; i < n; i += BLOCK) { vr += r[2 * i]; vi += r[2 * i + 1]; }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) { res[2 * blockIdx.x] = vr; res[2 * blockIdx.x + 1] = vi; }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_norm2(const cx<R>* __restrict__ a, uint64_t num,
                                                 double* __restrict__ part) {
    double acc = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        cx<R> v = a[j];
        acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    double dummy = 0;
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = acc; part[2 * blockIdx.x + 1] = 0; }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_inner(const cx<R>* __restrict__ a, const cx<R>* __restrict__ b,
                                                 uint64_t num, double* __restrict__ part) {
    double vr = 0, vi = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        cx<R> x = a[j], y = b[j];
        // conj(x) * y
        vr += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
        vi += (double)x.x * (double)y.y - (double)x.y * (double)y.x;
    }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = vr; part[2 * blockIdx.x + 1] = vi; }
}

// Pauli-string sums.  Workgroup (t, c) handles term t over chunk c of 2^N (DM)
// or 2^n (SV) indices and writes f_t * (its partial sum).
//  DM (reference distributed_densitymatrix.hpp:322-344 / misc.hpp:16-38): only
//  entries with column c = r ^ mxy are nonzero, each sigma product equals
//  (-i)^{#Y} (-1)^{par(c & myz)}, so Tr(P rho) = (-i)^{#Y} sum_r (-1)^{..} rho[r + 2^N c].
//  SV: <psi|P|psi> = i^{#Y} sum_j conj(psi_j) (-1)^{par(j1 & myz)} psi_{j1}, j1 = j ^ mxy.
constexpr uint64_t EXP_CHUNK = (uint64_t)BLOCK * 16;

template <typename R, bool DM>
__global__ __launch_bounds__(BLOCK) void k_pauli_expec(const cx<R>* __restrict__ a, unsigned nbits,
                                                       const PauliTermDev* __restrict__ terms,
                                                       unsigned chunks, double* __restrict__ part) {
    const unsigned t = blockIdx.x / chunks, c = blockIdx.x % chunks;
    const PauliTermDev term = terms[t];
    const uint64_t num = 1ULL << nbits;
    const uint64_t start = (uint64_t)c * EXP_CHUNK;
    const uint64_t end = start + EXP_CHUNK < num ? start + EXP_CHUNK : num;
    double vr = 0, vi = 0;
    for (uint64_t r = start + threadIdx.x; r < end; r += BLOCK) {
        uint64_t r1 = r ^ term.mxy;
        double sg = par64(r1 & term.myz) ? -1.0 : 1.0;
        if (DM) {
            cx<R> v = a[r + (r1 << nbits)];
            vr += sg * (double)v.x;
            vi += sg * (double)v.y;
        } else {
            cx<R> x = a[r], y = a[r1];
            vr += sg * ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
            vi += sg * ((double)x.x * (double)y.y - (double)x.y * (double)y.x);
        }
    }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = term.fr * vr - term.fi * vi;
        part[2 * blockIdx.x + 1] = term.fr * vi + term.fi * vr;
    }
}

// <a| f P |b> = f sum_j conj(a_j) (-1)^{par(j1 & myz)} b_{j1}, j1 = j ^ mxy; one partial per workgroup
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_pauli_element(const cx<R>* __restrict__ a, const cx<R>* __restrict__ b,
                                                         uint64_t num, PauliTermDev term,
                                                         double* __restrict__ part) {
    double vr = 0, vi = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        uint64_t j1 = j ^ term.mxy;
        double sg = par64(j1 & term.myz) ? -1.0 : 1.0;
        cx<R> x = a[j], y = b[j1];
        vr += sg * ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
        vi += sg * ((double)x.x * (double)y.y - (double)x.y * (double)y.x);
    }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = term.fr * vr - term.fi * vi;
        part[2 * blockIdx.x + 1] = term.fr * vi + term.fi * vr;
    }
}

// One step of the adjoint reverse sweep for a parametrised gadget, in one pass
// over phi and lambda: the term <lambda| f P |phi> (as k_pauli_element), then
// U^dag on both states with the arithmetic of k_pauli / k_pauli_diag / k_phase.
// Each pair (j, j ^ mxy) of both states is read once and written once.
// NT: streaming state access (ld_stream / st_stream) when every 128-byte line
// a wave touches is written whole by one store instruction (no X/Y bit inside
// a line)
template <typename R, bool NT>
__device__ __forceinline__ cx<R> adj_ld(const cx<R>* p) {
    if constexpr (NT) return ld_stream(p);
    else return *p;
}
template <typename R, bool NT>
__device__ __forceinline__ void adj_st(cx<R>* p, cx<R> v) {
    if constexpr (NT) st_stream(p, v);
    else *p = v;
}

// First block of a sweep kernel's grid-stride loop.  FP32 sweeps run 4-5%
// faster with each XCD on a contiguous eighth of the blocks (stream_block);
// FP64 ones 4-9% slower (24-qubit adjoint, same box), so they keep the
// round-robin order.
template <typename R>
__device__ __forceinline__ uint64_t adjoint_block() {
    if constexpr (sizeof(R) == 4) return stream_block();
    return blockIdx.x;
}

template <typename R, bool NT>
__global__ __launch_bounds__(BLOCK) void k_adjoint_step(cx<R>* __restrict__ phi, cx<R>* __restrict__ lam,
                                                        uint64_t numIts, unsigned hb, AdjointStepDev st,
                                                        double* __restrict__ rows, uint64_t stride) {
    const cx<R> th = mk<R>((R)st.thr, (R)st.thi), fe = mk<R>((R)st.fer, (R)st.fei);
    const cx<R> fo = mk<R>(-fe.x, -fe.y);
    const cx<R> e0 = mk<R>((R)st.e0r, (R)st.e0i), e1 = mk<R>((R)st.e1r, (R)st.e1i);
    double vr = 0, vi = 0;
    for (uint64_t k = adjoint_block<R>() * BLOCK + threadIdx.x; k < numIts; k += (uint64_t)gridDim.x * BLOCK) {
        if (st.mxy) {
            const uint64_t j0 = ins0(k, hb), j1 = j0 ^ st.mxy;
            const cx<R> p0 = adj_ld<R, NT>(phi + j0), p1 = adj_ld<R, NT>(phi + j1);
            const cx<R> l0 = adj_ld<R, NT>(lam + j0), l1 = adj_ld<R, NT>(lam + j1);
            const bool o0 = par64(j0 & st.myz) != 0, o1 = par64(j1 & st.myz) != 0;
            const double s0 = o0 ? -1.0 : 1.0, s1 = o1 ? -1.0 : 1.0;
            // conj(lambda_j) (-1)^{par(j' & myz)} phi_j' for j = j0 and j = j1
            vr += s1 * ((double)l0.x * (double)p1.x + (double)l0.y * (double)p1.y) +
                  s0 * ((double)l1.x * (double)p0.x + (double)l1.y * (double)p0.y);
            vi += s1 * ((double)l0.x * (double)p1.y - (double)l0.y * (double)p1.x) +
                  s0 * ((double)l1.x * (double)p0.y - (double)l1.y * (double)p0.x);
            const cx<R> f0 = o0 ? fo : fe, f1 = o1 ? fo : fe;
            adj_st<R, NT>(phi + j0, cfma(cmul(th, p0), f1, p1));
            adj_st<R, NT>(phi + j1, cfma(cmul(th, p1), f0, p0));
            adj_st<R, NT>(lam + j0, cfma(cmul(th, l0), f1, l1));
            adj_st<R, NT>(lam + j1, cfma(cmul(th, l1), f0, l0));
        } else {
            const cx<R> p = adj_ld<R, NT>(phi + k), l = adj_ld<R, NT>(lam + k);
            const bool odd = par64(k & st.myz) != 0;
            const double sg = odd ? -1.0 : 1.0;
            vr += sg * ((double)l.x * (double)p.x + (double)l.y * (double)p.y);
            vi += sg * ((double)l.x * (double)p.y - (double)l.y * (double)p.x);
            if (st.phase) {
                const cx<R> f = odd ? e1 : e0;
                adj_st<R, NT>(phi + k, cmul(p, f));
                adj_st<R, NT>(lam + k, cmul(l, f));
            } else {
                const cx<R> f = odd ? fo : fe;
                adj_st<R, NT>(phi + k, cfma(cmul(th, p), f, p));
                adj_st<R, NT>(lam + k, cfma(cmul(th, l), f, l));
            }
        }
    }
    block_sum2(vr, vi);
    if (threadIdx.x == 0) {
        double* row = rows + stride * st.slot + 2 * blockIdx.x;
        row[0] = st.fr * vr - st.fi * vi;
        row[1] = st.fr * vi + st.fi * vr;
    }
}

// A run of diagonal gadgets (AdjointDiagDev): conj(lambda_k) phi_k is formed
// once per amplitude and added, with gadget g's parity sign, to g's FP64
// accumulators; both states are multiplied by the product of the gadgets'
// U^dag factors, looked up by the amplitude's parity pattern (bit g = parity
// of gadget g) in a table each block builds in LDS.  Each accumulator then
// takes the same two-level reduction as k_adjoint_step: block sums here (one
// row per workgroup in the gadget's slot), then k_grad_sums.
template <typename R, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_adjoint_diag(cx<R>* __restrict__ phi, cx<R>* __restrict__ lam,
                                                        uint64_t num, AdjointDiagDev st, double* __restrict__ rows,
                                                        uint64_t stride) {
    __shared__ cx<R> tab[1u << ADJ_DIAG_MAX];
    const unsigned cnt = st.count;
    for (unsigned t = threadIdx.x; t < (1u << cnt); t += BLOCK) {
        cx<R> f = mk<R>((R)1, (R)0);
        for (unsigned g = 0; g < cnt; g++)
            f = cmul(f, ((t >> g) & 1u) ? mk<R>((R)st.e1r[g], (R)st.e1i[g]) : mk<R>((R)st.e0r[g], (R)st.e0i[g]));
        tab[t] = f;
    }
    __syncthreads();
    double vr[ADJ_DIAG_MAX], vi[ADJ_DIAG_MAX];
#pragma unroll
    for (unsigned g = 0; g < ADJ_DIAG_MAX; g++) { vr[g] = 0; vi[g] = 0; }
    for (uint64_t k = adjoint_block<R>() * BLOCK + threadIdx.x; k < num; k += (uint64_t)gridDim.x * BLOCK) {
        const cx<R> p = adj_ld<R, true>(phi + k), l = adj_ld<R, true>(lam + k);
        const double wr = (double)l.x * (double)p.x + (double)l.y * (double)p.y;
        const double wi = (double)l.x * (double)p.y - (double)l.y * (double)p.x;
        unsigned pat = 0;
#pragma unroll
        for (unsigned g = 0; g < ADJ_DIAG_MAX; g++) {
            if (g < cnt) {
                // WIDE: states of more than 32 qubits (64-bit parities)
                const unsigned odd = WIDE ? par64(k & st.myz[g]) : par32((uint32_t)k & (uint32_t)st.myz[g]);
                pat |= odd << g;
                vr[g] += odd ? -wr : wr;
                vi[g] += odd ? -wi : wi;
            }
        }
        const cx<R> f = tab[pat];
        adj_st<R, true>(phi + k, cmul(p, f));
        adj_st<R, true>(lam + k, cmul(l, f));
    }
#pragma unroll
    for (unsigned g = 0; g < ADJ_DIAG_MAX; g++) {
        if (g < cnt) {
            double a = vr[g], b = vi[g];
            block_sum2(a, b);
            if (threadIdx.x == 0) {
                double* row = rows + stride * st.slot[g] + 2 * blockIdx.x;
                row[0] = st.fr[g] * a - st.fi[g] * b;
                row[1] = st.fr[g] * b + st.fi[g] * a;
            }
            __syncthreads();    // block_sum2's wave partials are reused by the next gadget
        }
    }
}

// A run of G one-qubit X/Y gadgets on distinct qubits (AdjointQubitRunDev):
// each thread takes groups of 2^G amplitudes of phi and lambda (bit g of the
// group index = gadget g's qubit), adds every gadget's term from the loaded
// values (conj(lambda_j) s(j') phi_j' over its pairs j' = j ^ bit, as in
// k_adjoint_step), then applies every gadget's U^dag to its pairs and writes
// the group back.  Reduction as in k_adjoint_diag.
template <typename R, int G, bool NT>
__global__ __launch_bounds__(BLOCK) void k_adjoint_qrun(cx<R>* __restrict__ phi, cx<R>* __restrict__ lam,
                                                        uint64_t numGroups, AdjointQubitRunDev st,
                                                        double* __restrict__ rows, uint64_t stride) {
    constexpr unsigned E = 1u << G;
    double vr[G], vi[G];
#pragma unroll
    for (int g = 0; g < G; g++) { vr[g] = 0; vi[g] = 0; }
    uint64_t off[E];
#pragma unroll
    for (unsigned e = 0; e < E; e++) {
        off[e] = 0;
#pragma unroll
        for (int g = 0; g < G; g++) if ((e >> g) & 1u) off[e] |= 1ULL << st.q[g];
    }
    for (uint64_t t = adjoint_block<R>() * BLOCK + threadIdx.x; t < numGroups; t += (uint64_t)gridDim.x * BLOCK) {
        uint64_t x = t;
#pragma unroll
        for (int i = 0; i < G; i++) x = ins0(x, st.sorted[i]);
        cx<R> p[E], l[E];
#pragma unroll
        for (unsigned e = 0; e < E; e++) { p[e] = adj_ld<R, NT>(phi + (x | off[e])); l[e] = adj_ld<R, NT>(lam + (x | off[e])); }
#pragma unroll
        for (int g = 0; g < G; g++) {
            // s(j) = -1 where the gadget is Y and j has its qubit set (the pair's second element)
            const double s1 = st.y[g] ? -1.0 : 1.0;
#pragma unroll
            for (unsigned e = 0; e < E; e++) {
                if ((e >> g) & 1u) continue;
                const unsigned e1 = e | (1u << g);
                vr[g] += s1 * ((double)l[e].x * (double)p[e1].x + (double)l[e].y * (double)p[e1].y) +
                         ((double)l[e1].x * (double)p[e].x + (double)l[e1].y * (double)p[e].y);
                vi[g] += s1 * ((double)l[e].x * (double)p[e1].y - (double)l[e].y * (double)p[e1].x) +
                         ((double)l[e1].x * (double)p[e].y - (double)l[e1].y * (double)p[e].x);
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const cx<R> th = mk<R>((R)st.thr[g], (R)st.thi[g]), fe = mk<R>((R)st.fer[g], (R)st.fei[g]);
            const cx<R> f1 = st.y[g] ? mk<R>(-fe.x, -fe.y) : fe;    // factor of the element with the qubit set
#pragma unroll
            for (unsigned e = 0; e < E; e++) {
                if ((e >> g) & 1u) continue;
                const unsigned e1 = e | (1u << g);
                const cx<R> p0 = p[e], p1 = p[e1], l0 = l[e], l1 = l[e1];
                p[e] = cfma(cmul(th, p0), f1, p1);
                p[e1] = cfma(cmul(th, p1), fe, p0);
                l[e] = cfma(cmul(th, l0), f1, l1);
                l[e1] = cfma(cmul(th, l1), fe, l0);
            }
        }
#pragma unroll
        for (unsigned e = 0; e < E; e++) { adj_st<R, NT>(phi + (x | off[e]), p[e]); adj_st<R, NT>(lam + (x | off[e]), l[e]); }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        double a = vr[g], b = vi[g];
        block_sum2(a, b);
        if (threadIdx.x == 0) {
            double* row = rows + stride * st.slot[g] + 2 * blockIdx.x;
            row[0] = st.fr[g] * a - st.fi[g] * b;
            row[1] = st.fr[g] * b + st.fi[g] * a;
        }
        __syncthreads();    // block_sum2's wave partials are reused by the next gadget
    }
}

template <typename R>
__device__ __forceinline__ cx<R> shfl_xor_cx(cx<R> v, int m) {
    return mk<R>(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

// The same run when every qubit of it is below 6: lane l of a wave holds
// amplitude base + l of phi and lambda (whole 64-amplitude blocks, so every
// line is read whole), and the partner j ^ bit comes from lane l ^ bit.  Each
// element adds conj(lambda_j) s(j') phi_j' (j' = j ^ bit, s(j') = -1 for a Y
// gadget when j' has the bit set), which over a pair is k_adjoint_step's
// term; then each gadget in turn maps v_j <- th v_j + s(j') fe v_j'.
// DG: the diagonal run that follows in the sweep (dg, 1..ADJ_DIAG_MAX gadgets)
// is applied in the same pass, after the run, as k_adjoint_diag does it: its
// terms from the updated amplitudes, then its factors from an LDS table.
template <typename R, int G, bool DG, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_adjoint_qrun_lane(cx<R>* __restrict__ phi, cx<R>* __restrict__ lam,
                                                             uint64_t num, AdjointQubitRunDev st, AdjointDiagDev dg,
                                                             double* __restrict__ rows, uint64_t stride) {
    __shared__ cx<R> tab[DG ? 1u << ADJ_DIAG_MAX : 1u];
    const unsigned cnt = DG ? dg.count : 0u;
    if constexpr (DG) {
        for (unsigned t = threadIdx.x; t < (1u << cnt); t += BLOCK) {
            cx<R> f = mk<R>((R)1, (R)0);
            for (unsigned g = 0; g < cnt; g++)
                f = cmul(f, ((t >> g) & 1u) ? mk<R>((R)dg.e1r[g], (R)dg.e1i[g]) : mk<R>((R)dg.e0r[g], (R)dg.e0i[g]));
            tab[t] = f;
        }
        __syncthreads();
    }
    constexpr int DM = DG ? (int)ADJ_DIAG_MAX : 1;
    double vr[G], vi[G], dr[DM], di[DM];
#pragma unroll
    for (int g = 0; g < G; g++) { vr[g] = 0; vi[g] = 0; }
#pragma unroll
    for (int g = 0; g < DM; g++) { dr[g] = 0; di[g] = 0; }
    // num is a multiple of 64 and so is the stride: a wave's lanes stay together
    for (uint64_t j = adjoint_block<R>() * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        cx<R> p = adj_ld<R, true>(phi + j), l = adj_ld<R, true>(lam + j);
#pragma unroll
        for (int g = 0; g < G; g++) {
            const cx<R> pp = shfl_xor_cx(p, 1 << st.q[g]);
            const bool partnerSet = ((j >> st.q[g]) & 1ULL) == 0;
            const double sg = (st.y[g] && partnerSet) ? -1.0 : 1.0;
            vr[g] += sg * ((double)l.x * (double)pp.x + (double)l.y * (double)pp.y);
            vi[g] += sg * ((double)l.x * (double)pp.y - (double)l.y * (double)pp.x);
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int m = 1 << st.q[g];
            const cx<R> pp = shfl_xor_cx(p, m), lp = shfl_xor_cx(l, m);
            const bool partnerSet = ((j >> st.q[g]) & 1ULL) == 0;
            const cx<R> th = mk<R>((R)st.thr[g], (R)st.thi[g]);
            const cx<R> fe = mk<R>((R)st.fer[g], (R)st.fei[g]);
            const cx<R> f = (st.y[g] && partnerSet) ? mk<R>(-fe.x, -fe.y) : fe;
            p = cfma(cmul(th, p), f, pp);
            l = cfma(cmul(th, l), f, lp);
        }
        if constexpr (DG) {
            const double wr = (double)l.x * (double)p.x + (double)l.y * (double)p.y;
            const double wi = (double)l.x * (double)p.y - (double)l.y * (double)p.x;
            unsigned pat = 0;
#pragma unroll
            for (unsigned g = 0; g < ADJ_DIAG_MAX; g++) {
                if (g < cnt) {
                    // WIDE: states of more than 32 qubits (64-bit parities)
                    const unsigned odd = WIDE ? par64(j & dg.myz[g]) : par32((uint32_t)j & (uint32_t)dg.myz[g]);
                    pat |= odd << g;
                    dr[g] += odd ? -wr : wr;
                    di[g] += odd ? -wi : wi;
                }
            }
            const cx<R> fd = tab[pat];
            p = cmul(p, fd);
            l = cmul(l, fd);
        }
        adj_st<R, true>(phi + j, p);
        adj_st<R, true>(lam + j, l);
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        double a = vr[g], b = vi[g];
        block_sum2(a, b);
        if (threadIdx.x == 0) {
            double* row = rows + stride * st.slot[g] + 2 * blockIdx.x;
            row[0] = st.fr[g] * a - st.fi[g] * b;
            row[1] = st.fr[g] * b + st.fi[g] * a;
        }
        __syncthreads();    // block_sum2's wave partials are reused by the next gadget
    }
    if constexpr (DG) {
#pragma unroll
        for (unsigned g = 0; g < ADJ_DIAG_MAX; g++) {
            if (g < cnt) {
                double a = dr[g], b = di[g];
                block_sum2(a, b);
                if (threadIdx.x == 0) {
                    double* row = rows + stride * dg.slot[g] + 2 * blockIdx.x;
                    row[0] = dg.fr[g] * a - dg.fi[g] * b;
                    row[1] = dg.fr[g] * b + dg.fi[g] * a;
                }
                __syncthreads();
            }
        }
    }
}

// out[j] = sum_t f_t (-1)^{par(j1 & myz_t)} in[j1], j1 = j ^ mxy_t
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_apply_pauli_sum(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                           uint64_t num, const PauliTermDev* __restrict__ terms,
                                                           unsigned numTerms) {
    const uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= num) return;
    double ar = 0, ai = 0;
    for (unsigned t = 0; t < numTerms; t++) {
        const PauliTermDev term = terms[t];
        uint64_t j1 = j ^ term.mxy;
        double sg = par64(j1 & term.myz) ? -1.0 : 1.0;
        cx<R> v = in[j1];
        double fr = sg * term.fr, fi = sg * term.fi;
        ar += fr * (double)v.x - fi * (double)v.y;
        ai += fr * (double)v.y + fi * (double)v.x;
    }
    out[j] = mk<R>((R)ar, (R)ai);
}

struct TargetList { uint32_t t[32]; uint32_t k; };

// reference local_densitymatrix.hpp:134-164 (setBit defect corrected)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_partial_trace(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                         uint64_t numOut, unsigned N, uint64_t allMask,
                                                         TargetList tl) {
    const uint64_t l = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (l >= numOut) return;
    const uint64_t lm = ins0_mask(l, allMask);
    cx<R> acc = mk<R>(0, 0);
    const uint64_t numTraced = 1ULL << tl.k;
    for (uint64_t kk = 0; kk < numTraced; kk++) {
        uint64_t i = lm;
        for (unsigned b = 0; b < tl.k; b++)
            if ((kk >> b) & 1) i |= (1ULL << tl.t[b]) | (1ULL << (tl.t[b] + N));
        acc = cadd(acc, in[i]);
    }
    out[l] = acc;
}

// ---- initialisation ----------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit_open(uint64_t key, uint64_t c) {
    return ((double)(mix64(key + c) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_gaussian(cx<R>* __restrict__ a, uint64_t num, uint64_t key) {
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        double u1 = unit_open(key, 2 * j), u2 = unit_open(key, 2 * j + 1);
        double r = sqrt(-2.0 * log(u1));
        double ph = 6.283185307179586476925286766559 * u2;
        a[j] = mk<R>((R)(r * cos(ph)), (R)(r * sin(ph)));
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_scale_inv_sqrt(cx<R>* __restrict__ a, uint64_t num,
                                                          const double* __restrict__ norm2) {
    const double f = norm2[0] > 0 ? 1.0 / sqrt(norm2[0]) : 1.0;   // a zero vector stays zero
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK) {
        cx<R> v = a[j];
        a[j] = mk<R>((R)((double)v.x * f), (R)((double)v.y * f));
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_basis(cx<R>* __restrict__ a, uint64_t num, uint64_t index) {
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK)
        a[j] = mk<R>((R)(j == index ? 1 : 0), (R)0);
}

__global__ __launch_bounds__(BLOCK) void k_f64_to_f32(const cx<double>* __restrict__ in, cx<float>* __restrict__ out,
                                                      uint64_t num) {
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK)
        out[j] = mk<float>((float)in[j].x, (float)in[j].y);
}

__global__ __launch_bounds__(BLOCK) void k_f32_to_f64(const cx<float>* __restrict__ in, cx<double>* __restrict__ out,
                                                      uint64_t num) {
    for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < num; j += (uint64_t)gridDim.x * BLOCK)
        out[j] = mk<double>((double)in[j].x, (double)in[j].y);
}

unsigned grid_stride_blocks(uint64_t num) {
    uint64_t b = (num + BLOCK - 1) / BLOCK;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

}  // namespace

uint64_t reduce_workspace_doubles(uint64_t num) {
    (void)num;
    return 2ULL * RED_MAX_BLOCKS;
}

template <typename R>
void launch_norm2(hipStream_t s, const cx<R>* a, uint64_t num, double* ws, double* dout) {
    unsigned nb = red_blocks(num);
    hipLaunchKernelGGL(k_norm2<R>, dim3(nb), dim3(BLOCK), 0, s, a, num, ws);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLOCK), 0, s, ws, nb, dout);
}

template <typename R>
void launch_inner(hipStream_t s, const cx<R>* a, const cx<R>* b, uint64_t num, double* ws, double* dout) {
    unsigned nb = red_blocks(num);
    hipLaunchKernelGGL(k_inner<R>, dim3(nb), dim3(BLOCK), 0, s, a, b, num, ws);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLOCK), 0, s, ws, nb, dout);
}

template <typename R>
void launch_pauli_element(hipStream_t s, const cx<R>* a, const cx<R>* b, uint64_t num, PauliTermDev term,
                          double* ws, double* dout) {
    unsigned nb = red_blocks(num);
    hipLaunchKernelGGL(k_pauli_element<R>, dim3(nb), dim3(BLOCK), 0, s, a, b, num, term, ws);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLOCK), 0, s, ws, nb, dout);
}

uint64_t adjoint_row_stride(unsigned nq) { return 2ULL * red_blocks(1ULL << nq); }

void launch_grad_sums(hipStream_t s, const double* rows, uint64_t stride, const uint32_t* nb, unsigned numSlots,
                      double* res) {
    if (numSlots) hipLaunchKernelGGL(k_grad_sums, dim3(numSlots), dim3(BLOCK), 0, s, rows, stride, nb, res);
}

template <typename R>
unsigned launch_adjoint_step(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointStepDev& st,
                             double* rows, uint64_t stride) {
    const uint64_t numIts = st.mxy ? 1ULL << (nq - 1) : 1ULL << nq;
    const unsigned hb = st.mxy ? 63u - (unsigned)__builtin_clzll(st.mxy) : 0u;
    unsigned nb = red_blocks(numIts);
    const uint64_t lineBits = 128 / sizeof(cx<R>) - 1;
    auto* kern = (st.mxy & lineBits) ? k_adjoint_step<R, false> : k_adjoint_step<R, true>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, s, phi, lam, numIts, hb, st, rows, stride);
    return nb;
}

template <typename R>
unsigned launch_adjoint_diag(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointDiagDev& st,
                             double* rows, uint64_t stride) {
    const uint64_t num = 1ULL << nq;
    const unsigned nb = red_blocks(num);
    auto* kern = nq > 32 ? k_adjoint_diag<R, true> : k_adjoint_diag<R, false>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, s, phi, lam, num, st, rows, stride);
    return nb;
}

template <typename R>
unsigned launch_adjoint_qrun(hipStream_t s, cx<R>* phi, cx<R>* lam, unsigned nq, const AdjointQubitRunDev& st,
                             double* rows, uint64_t stride, const AdjointDiagDev* tail) {
    const unsigned G = st.count;
    if (G < 2 || G > ADJ_QRUN_MAX || nq < G) fail(FSA_ERR_INTERNAL, "adjoint qubit run of unsupported length");
    const uint64_t numGroups = 1ULL << (nq - G);
    uint64_t bits = 0;
    for (unsigned g = 0; g < G; g++) bits |= 1ULL << st.q[g];
    if (bits < 64 && nq >= 6) {
        // every qubit of the run inside a wave's 64 consecutive amplitudes
        const uint64_t num = 1ULL << nq;
        const unsigned nbl = red_blocks(num);
        AdjointDiagDev dg;
        std::memset(&dg, 0, sizeof(dg));
        if (tail) dg = *tail;
#define FSA_QL(GV, DGV, WV)                                                                                       \
        hipLaunchKernelGGL((k_adjoint_qrun_lane<R, GV, DGV, WV>), dim3(nbl), dim3(BLOCK), 0, s, phi, lam, num, st, \
                           dg, rows, stride)
        const bool wide = nq > 32;
        if (G == 2) {
            if (!tail) FSA_QL(2, false, false);
            else if (wide) FSA_QL(2, true, true);
            else FSA_QL(2, true, false);
        } else {
            if (!tail) FSA_QL(3, false, false);
            else if (wide) FSA_QL(3, true, true);
            else FSA_QL(3, true, false);
        }
#undef FSA_QL
        return nbl;
    }
    if (tail) fail(FSA_ERR_INTERNAL, "a diagonal tail needs a qubit run below qubit 6");
    const unsigned nb = red_blocks(numGroups);
    const bool nt = (bits & (128 / sizeof(cx<R>) - 1)) == 0;
#define FSA_QR(GV)                                                                                                \
    if (G == GV) {                                                                                                \
        auto* kern = nt ? k_adjoint_qrun<R, GV, true> : k_adjoint_qrun<R, GV, false>;                             \
        hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, s, phi, lam, numGroups, st, rows, stride);             \
    }
    FSA_QR(2) FSA_QR(3)
#undef FSA_QR
    return nb;
}

// ws must hold 2 * numTerms * ceil(2^nbits / EXP_CHUNK) doubles
template <typename R>
void launch_dm_expec(hipStream_t s, const cx<R>* rho, unsigned N, const PauliTermDev* dterms,
                     unsigned numTerms, double* ws, double* dout) {
    unsigned chunks = (unsigned)(((1ULL << N) + EXP_CHUNK - 1) / EXP_CHUNK);
    unsigned nb = numTerms * chunks;
    hipLaunchKernelGGL((k_pauli_expec<R, true>), dim3(nb), dim3(BLOCK), 0, s, rho, N, dterms, chunks, ws);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLOCK), 0, s, ws, nb, dout);
}

template <typename R>
void launch_sv_expec(hipStream_t s, const cx<R>* psi, unsigned n, const PauliTermDev* dterms,
                     unsigned numTerms, double* ws, double* dout) {
    unsigned chunks = (unsigned)(((1ULL << n) + EXP_CHUNK - 1) / EXP_CHUNK);
    unsigned nb = numTerms * chunks;
    hipLaunchKernelGGL((k_pauli_expec<R, false>), dim3(nb), dim3(BLOCK), 0, s, psi, n, dterms, chunks, ws);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLOCK), 0, s, ws, nb, dout);
}

uint64_t expec_workspace_doubles(unsigned nbits, unsigned numTerms) {
    return 2ULL * numTerms * ((((1ULL << nbits) + EXP_CHUNK - 1) / EXP_CHUNK));
}

template <typename R>
void launch_apply_pauli_sum(hipStream_t s, const cx<R>* in, cx<R>* out, unsigned n, const PauliTermDev* dterms,
                            unsigned numTerms) {
    uint64_t num = 1ULL << n;
    hipLaunchKernelGGL(k_apply_pauli_sum<R>, dim3((unsigned)((num + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, in,
                       out, num, dterms, numTerms);
}

template <typename R>
void launch_partial_trace(hipStream_t s, const cx<R>* in, cx<R>* out, unsigned N, const unsigned* sortedTargets,
                          unsigned k) {
    TargetList tl{};
    uint64_t allMask = 0;
    for (unsigned i = 0; i < k; i++) {
        tl.t[i] = sortedTargets[i];
        allMask |= (1ULL << sortedTargets[i]) | (1ULL << (sortedTargets[i] + N));
    }
    tl.k = k;
    uint64_t numOut = 1ULL << (2 * (N - k));
    hipLaunchKernelGGL(k_partial_trace<R>, dim3((unsigned)((numOut + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, in,
                       out, numOut, N, allMask, tl);
}

template <typename R>
void launch_init_gaussian(hipStream_t s, cx<R>* a, unsigned nq, uint64_t seed) {
    // the host mixes the seed the same way the oracle does (key = mix64(seed))
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    uint64_t key = z ^ (z >> 31);
    uint64_t num = 1ULL << nq;
    hipLaunchKernelGGL(k_gaussian<R>, dim3(grid_stride_blocks(num)), dim3(BLOCK), 0, s, a, num, key);
}

template <typename R>
void launch_scale_by_inv_sqrt(hipStream_t s, cx<R>* a, uint64_t num, const double* dnorm2) {
    hipLaunchKernelGGL(k_scale_inv_sqrt<R>, dim3(grid_stride_blocks(num)), dim3(BLOCK), 0, s, a, num, dnorm2);
}

template <typename R>
void launch_init_basis(hipStream_t s, cx<R>* a, uint64_t num, uint64_t index) {
    hipLaunchKernelGGL(k_basis<R>, dim3(grid_stride_blocks(num)), dim3(BLOCK), 0, s, a, num, index);
}

void launch_f64_to_f32(hipStream_t s, const cx<double>* in, cx<float>* out, uint64_t num) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_stride_blocks(num)), dim3(BLOCK), 0, s, in, out, num);
}

void launch_f32_to_f64(hipStream_t s, const cx<float>* in, cx<double>* out, uint64_t num) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_stride_blocks(num)), dim3(BLOCK), 0, s, in, out, num);
}

#define FSA_INSTANTIATE_RED(R)                                                                               \
    template void launch_norm2<R>(hipStream_t, const cx<R>*, uint64_t, double*, double*);                     \
    template void launch_inner<R>(hipStream_t, const cx<R>*, const cx<R>*, uint64_t, double*, double*);       \
    template unsigned launch_adjoint_step<R>(hipStream_t, cx<R>*, cx<R>*, unsigned, const AdjointStepDev&,       \
                                             double*, uint64_t);                                                \
    template unsigned launch_adjoint_diag<R>(hipStream_t, cx<R>*, cx<R>*, unsigned, const AdjointDiagDev&,       \
                                             double*, uint64_t);                                                \
    template unsigned launch_adjoint_qrun<R>(hipStream_t, cx<R>*, cx<R>*, unsigned, const AdjointQubitRunDev&,   \
                                             double*, uint64_t, const AdjointDiagDev*);                         \
    template void launch_pauli_element<R>(hipStream_t, const cx<R>*, const cx<R>*, uint64_t, PauliTermDev,     \
                                          double*, double*);                                                  \
    template void launch_dm_expec<R>(hipStream_t, const cx<R>*, unsigned, const PauliTermDev*, unsigned,       \
                                     double*, double*);                                                       \
    template void launch_sv_expec<R>(hipStream_t, const cx<R>*, unsigned, const PauliTermDev*, unsigned,       \
                                     double*, double*);                                                       \
    template void launch_apply_pauli_sum<R>(hipStream_t, const cx<R>*, cx<R>*, unsigned, const PauliTermDev*, \
                                            unsigned);                                                        \
    template void launch_partial_trace<R>(hipStream_t, const cx<R>*, cx<R>*, unsigned, const unsigned*,       \
                                          unsigned);                                                          \
    template void launch_init_gaussian<R>(hipStream_t, cx<R>*, unsigned, uint64_t);                           \
    template void launch_scale_by_inv_sqrt<R>(hipStream_t, cx<R>*, uint64_t, const double*);                  \
    template void launch_init_basis<R>(hipStream_t, cx<R>*, uint64_t, uint64_t);

FSA_INSTANTIATE_RED(double)
FSA_INSTANTIATE_RED(float)

}  // namespace fsa
