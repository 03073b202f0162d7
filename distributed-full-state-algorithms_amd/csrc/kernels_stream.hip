// kernels_stream.hip -- one-pass streaming kernels, one launch per operation.
//
// Each kernel replaces one OpenMP loop of the reference's local_* functions
// (cited per kernel).  Threads own independent amplitude pairs/quads/groups;
// every thread issues all of its loads before any arithmetic (ITEMS items per
// thread) so each wave keeps several 16-byte loads in flight.  Consecutive
// threads own consecutive items, so for any target position a wave touches
// whole cache lines.
#include "launch.hpp"

namespace fsa {

namespace {

constexpr int ITEMS = 4;

// Amplitude access: the streaming form (ld_stream / st_stream) when every
// 128-byte line a wave touches is written whole by one store instruction
// (NT); kernels whose mixed bits fall inside a line write alternating
// elements per instruction and stay on plain cached accesses.
template <bool NT, typename R>
__device__ __forceinline__ cx<R> ld(const cx<R>* p) {
    if constexpr (NT) return ld_stream(p);
    else return *p;
}
template <bool NT, typename R>
__device__ __forceinline__ void st(cx<R>* p, cx<R> v) {
    if constexpr (NT) st_stream(p, v);
    else *p = v;
}
// amplitude-index bits inside one 128-byte line
template <typename R>
constexpr uint64_t line_bits() { return 128 / sizeof(cx<R>) - 1; }

inline unsigned blocks_for(uint64_t numItems, int itemsPerThread) {
    uint64_t per = (uint64_t)BLOCK * itemsPerThread;
    uint64_t b = (numItems + per - 1) / per;
    return (unsigned)(b ? b : 1);
}

// local_statevector.hpp:14-51 (oneTargGate and manyCtrlOneTargGate): pair
// (i0, i1 = i0 | 2^t) with every control bit set.  qmask = controls | target.
template <typename R, bool NT>
__global__ __launch_bounds__(BLOCK) void k_mat1(cx<R>* __restrict__ a, uint64_t numIts, unsigned t,
                                                uint64_t qmask, uint64_t cmask, Mat2<R> g) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const uint64_t tb = 1ULL << t;
    cx<R> v0[ITEMS], v1[ITEMS];
    uint64_t idx[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        idx[p] = ins0_mask(j, qmask) | cmask;
        if (j < numIts) { v0[p] = ld<NT>(a + idx[p]); v1[p] = ld<NT>(a + (idx[p] | tb)); }
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < numIts) {
            st<NT>(a + idx[p], cfma(cmul(g.m[0], v0[p]), g.m[1], v1[p]));
            st<NT>(a + (idx[p] | tb), cfma(cmul(g.m[2], v0[p]), g.m[3], v1[p]));
        }
    }
}

// The same gate, uncontrolled, on a target inside a 128-byte line: lane l of
// a wave holds amplitude base + l (whole lines, non-temporal for FP64) and
// takes its partner from lane l ^ 2^t; the arithmetic is k_mat1's.  Needs the
// state to hold whole waves of amplitudes (2^n >= 64).  At 30 qubits FP64,
// target 0: 6.07 ms against 6.63 ms for k_mat1's cached form.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_mat1_lane(cx<R>* __restrict__ a, uint64_t num, unsigned t, Mat2<R> g) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const int m = 1 << t;
    cx<R> v[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        const uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) v[p] = ld<true>(a + j);
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        const uint64_t j = first + (uint64_t)p * BLOCK;
        // whole waves share j < num (num and BLOCK are multiples of 64)
        if (j < num) {
            const cx<R> w = mk<R>(__shfl_xor(v[p].x, m, 64), __shfl_xor(v[p].y, m, 64));
            const cx<R> r = ((j >> t) & 1ULL) ? cfma(cmul(g.m[2], w), g.m[3], v[p]) : cfma(cmul(g.m[0], v[p]), g.m[1], w);
            st<true>(a + j, r);
        }
    }
}

// local_statevector.hpp:54-69: swap amplitudes |..1..0..> <-> |..0..1..>
template <typename R, bool NT>
__global__ __launch_bounds__(BLOCK) void k_swap(cx<R>* __restrict__ a, uint64_t numIts, uint64_t b1,
                                                uint64_t b2) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const uint64_t m = b1 | b2;
    R ax[ITEMS], ay[ITEMS], bx[ITEMS], by[ITEMS];    // split re/im keeps them in registers
    uint64_t idx[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        idx[p] = ins0_mask(k, m);
        if (k < numIts) {
            cx<R> u = ld<NT>(a + (idx[p] | b1)), v = ld<NT>(a + (idx[p] | b2));
            ax[p] = u.x; ay[p] = u.y; bx[p] = v.x; by[p] = v.y;
        }
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        if (k < numIts) { st<NT>(a + (idx[p] | b1), mk<R>(bx[p], by[p])); st<NT>(a + (idx[p] | b2), mk<R>(ax[p], ay[p])); }
    }
}

// local_statevector.hpp:102-135 (with the setBit defect corrected): pairs
// (j0, j1 = j0 ^ mxy).  new[j0] = this*a0 + f(j1)*a1, new[j1] = this*a1 + f(j0)*a0,
// f(j) = evenFac * (-1)^{par(j & myz)}, evenFac = other * i^{#Y}.
template <typename R, bool NT>
__global__ __launch_bounds__(BLOCK) void k_pauli(cx<R>* __restrict__ a, uint64_t numIts, unsigned hb,
                                                 uint64_t mxy, uint64_t myz, cx<R> th, cx<R> fe) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const cx<R> fo = mk<R>(-fe.x, -fe.y);
    cx<R> v0[ITEMS], v1[ITEMS];
    uint64_t idx[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        idx[p] = ins0(k, hb);
        if (k < numIts) { v0[p] = ld<NT>(a + idx[p]); v1[p] = ld<NT>(a + (idx[p] ^ mxy)); }
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        if (k < numIts) {
            uint64_t j0 = idx[p], j1 = j0 ^ mxy;
            cx<R> f0 = par64(j0 & myz) ? fo : fe;
            cx<R> f1 = par64(j1 & myz) ? fo : fe;
            st<NT>(a + j0, cfma(cmul(th, v0[p]), f1, v1[p]));
            st<NT>(a + j1, cfma(cmul(th, v1[p]), f0, v0[p]));
        }
    }
}

// all-Z Pauli string: the pair degenerates to one amplitude (D2 contract)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_pauli_diag(cx<R>* __restrict__ a, uint64_t num, uint64_t myz,
                                                      cx<R> th, cx<R> fe) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const cx<R> fo = mk<R>(-fe.x, -fe.y);
    cx<R> v[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) v[p] = ld_stream(a + j);
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) {
            cx<R> f = par64(j & myz) ? fo : fe;
            st_stream(a + j, cfma(cmul(th, v[p]), f, v[p]));
        }
    }
}

// local_statevector.hpp:138-153
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_phase(cx<R>* __restrict__ a, uint64_t num, uint64_t mask,
                                                 cx<R> f0, cx<R> f1) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    cx<R> v[ITEMS];
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) v[p] = ld_stream(a + j);
    }
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) st_stream(a + j, cmul(v[p], par64(j & mask) ? f1 : f0));
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_scale(cx<R>* __restrict__ a, uint64_t num, R f) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) st_stream(a + j, cscale(ld_stream(a + j), f));
    }
}

// local_statevector.hpp:72-99 (setBit defect corrected): one thread per group
// of 2^K amplitudes, matrix staged in LDS and read with wave-uniform addresses.
template <typename R, int K, bool NT>
__global__ __launch_bounds__(BLOCK) void k_matk_reg(cx<R>* __restrict__ a, uint64_t numGroups,
                                                    uint64_t tmask, Offs32<R> off,
                                                    const cx<R>* __restrict__ mat) {
    constexpr int D = 1 << K;
    __shared__ cx<R> sm[D * D];
    for (int i = threadIdx.x; i < D * D; i += BLOCK) sm[i] = mat[i];
    __syncthreads();
    const uint64_t g = stream_block() * BLOCK + threadIdx.x;
    if (g >= numGroups) return;
    const uint64_t base = ins0_mask(g, tmask);
    cx<R> v[D];
#pragma unroll
    for (int l = 0; l < D; l++) v[l] = ld<NT>(a + (base | off.o[l]));
#pragma unroll
    for (int r = 0; r < D; r++) {
        cx<R> acc = mk<R>(0, 0);
#pragma unroll
        for (int l = 0; l < D; l++) acc = cfma(acc, sm[r * D + l], v[l]);
        st<NT>(a + (base | off.o[r]), acc);
    }
}

// dense gate with 6 <= k <= 12 targets: one workgroup per group, the group in
// LDS, rows split over threads, transposed matrix so row reads coalesce.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_matk_lds(cx<R>* __restrict__ a, unsigned k, uint64_t tmask,
                                                    const uint64_t* __restrict__ offs,
                                                    const cx<R>* __restrict__ matT) {
    extern __shared__ unsigned char smem_raw[];
    cx<R>* sv = reinterpret_cast<cx<R>*>(smem_raw);
    const uint32_t D = 1u << k;
    const uint64_t base = ins0_mask((uint64_t)blockIdx.x, tmask);
    for (uint32_t l = threadIdx.x; l < D; l += BLOCK) sv[l] = a[base | offs[l]];
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < D; r += BLOCK) {
        cx<R> acc = mk<R>(0, 0);
        for (uint32_t l = 0; l < D; l++) acc = cfma(acc, matT[(uint64_t)l * D + r], sv[l]);
        a[base | offs[r]] = acc;
    }
}

// ---- density-matrix channels (local_densitymatrix.hpp) -------------------

// :28-41 -- scale rho_{01}, rho_{10} of qubit q
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_dephase1(cx<R>* __restrict__ a, uint64_t numIts, unsigned q,
                                                    unsigned alt, R fac) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const uint64_t m = (1ULL << q) | (1ULL << alt);
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        if (k < numIts) {
            uint64_t j = ins0_mask(k, m);
            uint64_t j01 = j | (1ULL << q), j10 = j | (1ULL << alt);
            cx<R> x = a[j01], y = a[j10];
            a[j01] = cscale(x, fac);
            a[j10] = cscale(y, fac);
        }
    }
}

// :45-60 -- scale elements whose row and column differ on q1 or on q2
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_dephase2(cx<R>* __restrict__ a, uint64_t num, unsigned q1,
                                                    unsigned a1, unsigned q2, unsigned a2, R fac) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t j = first + (uint64_t)p * BLOCK;
        if (j < num) {
            uint64_t f = (((j >> q1) ^ (j >> a1)) | ((j >> q2) ^ (j >> a2))) & 1ULL;
            if (f) a[j] = cscale(a[j], fac);
        }
    }
}

// :63-81 -- mix rho_00/rho_11 of qubit q, scale the coherences
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_depol1(cx<R>* __restrict__ a, uint64_t numIts, unsigned q,
                                                  unsigned alt, R c1, R c2, R c3) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const uint64_t bq = 1ULL << q, ba = 1ULL << alt;
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        if (k < numIts) {
            uint64_t j00 = ins0_mask(k, bq | ba);
            cx<R> v00 = a[j00], v01 = a[j00 | bq], v10 = a[j00 | ba], v11 = a[j00 | bq | ba];
            a[j00] = mk<R>(c2 * v00.x + c1 * v11.x, c2 * v00.y + c1 * v11.y);
            a[j00 | bq] = cscale(v01, c3);
            a[j00 | ba] = cscale(v10, c3);
            a[j00 | bq | ba] = mk<R>(c1 * v00.x + c2 * v11.x, c1 * v00.y + c2 * v11.y);
        }
    }
}

// :83-108 in ONE pass: per group of 16 (bits q0..q3; the channel passes q1, q2,
// q1+N, q2+N) scale the 12
// off-block-diagonal entries by 1 + c3 and mix the 4 block-diagonal ones as
// c1*x + c2*sum.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_depol2(cx<R>* __restrict__ a, uint64_t numIts, uint64_t mask,
                                                  uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3,
                                                  R c1, R c2, R c3) {
    const uint64_t k = stream_block() * BLOCK + threadIdx.x;
    if (k >= numIts) return;
    const uint64_t base = ins0_mask(k, mask);
    const uint64_t bit[4] = {b0, b1, b2, b3};   // q1, q2, q1+N, q2+N
    cx<R> v[16];
#pragma unroll
    for (int l = 0; l < 16; l++) {
        uint64_t i = base;
#pragma unroll
        for (int b = 0; b < 4; b++) if ((l >> b) & 1) i |= bit[b];
        v[l] = a[i];
    }
    // block-diagonal: row bits == column bits on both qubits: l in {0, 5, 10, 15}
    cx<R> sum = cadd(cadd(v[0], v[5]), cadd(v[10], v[15]));
    const R s = (R)1 + c3;
#pragma unroll
    for (int l = 0; l < 16; l++) {
        uint64_t i = base;
#pragma unroll
        for (int b = 0; b < 4; b++) if ((l >> b) & 1) i |= bit[b];
        bool diag = (l == 0 || l == 5 || l == 10 || l == 15);
        cx<R> r = diag ? mk<R>(c1 * v[l].x + c2 * sum.x, c1 * v[l].y + c2 * sum.y) : cscale(v[l], s);
        a[i] = r;
    }
}

// :111-131
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_damp(cx<R>* __restrict__ a, uint64_t numIts, unsigned q,
                                                unsigned alt, R prob, R c1, R c2) {
    const uint64_t first = stream_block() * (BLOCK * ITEMS) + threadIdx.x;
    const uint64_t bq = 1ULL << q, ba = 1ULL << alt;
#pragma unroll
    for (int p = 0; p < ITEMS; p++) {
        uint64_t k = first + (uint64_t)p * BLOCK;
        if (k < numIts) {
            uint64_t j00 = ins0_mask(k, bq | ba);
            cx<R> v00 = a[j00], v01 = a[j00 | bq], v10 = a[j00 | ba], v11 = a[j00 | bq | ba];
            a[j00] = mk<R>(v00.x + prob * v11.x, v00.y + prob * v11.y);
            a[j00 | bq] = cscale(v01, c1);
            a[j00 | ba] = cscale(v10, c1);
            a[j00 | bq | ba] = cscale(v11, c2);
        }
    }
}

}  // namespace

// ---- launchers -------------------------------------------------------------

template <typename R>
void launch_mat1(hipStream_t s, cx<R>* a, unsigned nq, unsigned t, uint64_t ctrlMask, const Mat2<R>& g) {
    unsigned nc = (unsigned)__builtin_popcountll(ctrlMask);
    uint64_t numIts = (1ULL << nq) >> (nc + 1);
    const uint64_t qmask = ctrlMask | (1ULL << t);
    if (ctrlMask == 0 && ((1ULL << t) & line_bits<R>()) && nq >= 6) {
        const uint64_t num = 1ULL << nq;
        hipLaunchKernelGGL(k_mat1_lane<R>, dim3(blocks_for(num, ITEMS)), dim3(BLOCK), 0, s, a, num, t, g);
        return;
    }
    auto* kern = (qmask & line_bits<R>()) ? k_mat1<R, false> : k_mat1<R, true>;
    hipLaunchKernelGGL(kern, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, t, qmask, ctrlMask, g);
}

template <typename R>
void launch_swap(hipStream_t s, cx<R>* a, unsigned nq, unsigned q1, unsigned q2) {
    uint64_t numIts = (1ULL << nq) >> 2;
    auto* kern = (((1ULL << q1) | (1ULL << q2)) & line_bits<R>()) ? k_swap<R, false> : k_swap<R, true>;
    hipLaunchKernelGGL(kern, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, 1ULL << q1, 1ULL << q2);
}

template <typename R>
void launch_pauli(hipStream_t s, cx<R>* a, unsigned nq, uint64_t mxy, uint64_t myz, cx<R> th, cx<R> fe) {
    uint64_t num = 1ULL << nq;
    if (mxy == 0) {
        hipLaunchKernelGGL(k_pauli_diag<R>, dim3(blocks_for(num, ITEMS)), dim3(BLOCK), 0, s, a, num, myz, th, fe);
        return;
    }
    unsigned hb = 63u - (unsigned)__builtin_clzll(mxy);
    uint64_t numIts = num >> 1;
    auto* kern = (mxy & line_bits<R>()) ? k_pauli<R, false> : k_pauli<R, true>;
    hipLaunchKernelGGL(kern, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, hb, mxy, myz, th, fe);
}

template <typename R>
void launch_phase(hipStream_t s, cx<R>* a, unsigned nq, uint64_t mask, cx<R> f0, cx<R> f1) {
    uint64_t num = 1ULL << nq;
    hipLaunchKernelGGL(k_phase<R>, dim3(blocks_for(num, ITEMS)), dim3(BLOCK), 0, s, a, num, mask, f0, f1);
}

template <typename R>
void launch_scale(hipStream_t s, cx<R>* a, uint64_t num, R f) {
    hipLaunchKernelGGL(k_scale<R>, dim3(blocks_for(num, ITEMS)), dim3(BLOCK), 0, s, a, num, f);
}

template <typename R>
void launch_matk(hipStream_t s, cx<R>* a, unsigned nq, const unsigned* targets, unsigned k,
                 const cx<R>* dmat, const cx<R>* dmatT) {
    uint64_t tmask = 0;
    for (unsigned i = 0; i < k; i++) tmask |= 1ULL << targets[i];
    uint64_t numGroups = (1ULL << nq) >> k;
    if (k <= 5) {
        Offs32<R> off{};
        for (unsigned l = 0; l < (1u << k); l++) {
            uint64_t o = 0;
            for (unsigned i = 0; i < k; i++) if ((l >> i) & 1) o |= 1ULL << targets[i];
            off.o[l] = o;
        }
        dim3 grid((unsigned)((numGroups + BLOCK - 1) / BLOCK));
        const bool nt = (tmask & line_bits<R>()) == 0;
#define FSA_MATK(K) hipLaunchKernelGGL((nt ? k_matk_reg<R, K, true> : k_matk_reg<R, K, false>), grid, dim3(BLOCK), 0, s, \
                                       a, numGroups, tmask, off, dmat)
        switch (k) {
        case 1: FSA_MATK(1); break;
        case 2: FSA_MATK(2); break;
        case 3: FSA_MATK(3); break;
        case 4: FSA_MATK(4); break;
        default: FSA_MATK(5); break;
        }
#undef FSA_MATK
        return;
    }
    // offsets live after the transposed matrix in the same parameter block
    const uint64_t* offs = reinterpret_cast<const uint64_t*>(dmatT + (1ULL << (2 * k)));
    size_t lds = sizeof(cx<R>) << k;
    hipLaunchKernelGGL(k_matk_lds<R>, dim3((unsigned)numGroups), dim3(BLOCK), lds, s, a, k, tmask, offs, dmatT);
}

template <typename R>
void launch_dephase1(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R fac) {
    uint64_t numIts = (1ULL << (2 * N)) >> 2;
    hipLaunchKernelGGL(k_dephase1<R>, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, q, q + N, fac);
}

template <typename R>
void launch_dephase2(hipStream_t s, cx<R>* a, unsigned N, unsigned q1, unsigned q2, R term) {
    uint64_t num = 1ULL << (2 * N);
    hipLaunchKernelGGL(k_dephase2<R>, dim3(blocks_for(num, ITEMS)), dim3(BLOCK), 0, s, a, num, q1, q1 + N,
                       q2, q2 + N, (R)1 + term);
}

template <typename R>
void launch_depol1(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R c1, R c2, R c3) {
    uint64_t numIts = (1ULL << (2 * N)) >> 2;
    hipLaunchKernelGGL(k_depol1<R>, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, q, q + N,
                       c1, c2, c3);
}

template <typename R>
void launch_depol2(hipStream_t s, cx<R>* a, unsigned nq, const unsigned* q, R c1, R c2, R c3) {
    uint64_t numIts = (1ULL << nq) >> 4;
    uint64_t b0 = 1ULL << q[0], b1 = 1ULL << q[1], b2 = 1ULL << q[2], b3 = 1ULL << q[3];
    hipLaunchKernelGGL(k_depol2<R>, dim3((unsigned)((numIts + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, a,
                       numIts, b0 | b1 | b2 | b3, b0, b1, b2, b3, c1, c2, c3);
}

template <typename R>
void launch_damp(hipStream_t s, cx<R>* a, unsigned N, unsigned q, R prob, R c1, R c2) {
    uint64_t numIts = (1ULL << (2 * N)) >> 2;
    hipLaunchKernelGGL(k_damp<R>, dim3(blocks_for(numIts, ITEMS)), dim3(BLOCK), 0, s, a, numIts, q, q + N,
                       prob, c1, c2);
}

#define FSA_INSTANTIATE(R)                                                                              \
    template void launch_mat1<R>(hipStream_t, cx<R>*, unsigned, unsigned, uint64_t, const Mat2<R>&);   \
    template void launch_swap<R>(hipStream_t, cx<R>*, unsigned, unsigned, unsigned);                   \
    template void launch_matk<R>(hipStream_t, cx<R>*, unsigned, const unsigned*, unsigned,             \
                                 const cx<R>*, const cx<R>*);                                         \
    template void launch_pauli<R>(hipStream_t, cx<R>*, unsigned, uint64_t, uint64_t, cx<R>, cx<R>);    \
    template void launch_phase<R>(hipStream_t, cx<R>*, unsigned, uint64_t, cx<R>, cx<R>);              \
    template void launch_scale<R>(hipStream_t, cx<R>*, uint64_t, R);                                   \
    template void launch_dephase1<R>(hipStream_t, cx<R>*, unsigned, unsigned, R);                      \
    template void launch_dephase2<R>(hipStream_t, cx<R>*, unsigned, unsigned, unsigned, R);            \
    template void launch_depol1<R>(hipStream_t, cx<R>*, unsigned, unsigned, R, R, R);                  \
    template void launch_depol2<R>(hipStream_t, cx<R>*, unsigned, const unsigned*, R, R, R);             \
    template void launch_damp<R>(hipStream_t, cx<R>*, unsigned, unsigned, R, R, R);

FSA_INSTANTIATE(double)
FSA_INSTANTIATE(float)

}  // namespace fsa
