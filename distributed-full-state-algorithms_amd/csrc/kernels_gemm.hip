// kernels_gemm.hip -- a dense gate of 6 <= k <= 12 targets as one complex GEMM
// on MFMA (local_statevector.hpp:72-99 with the setBit defect corrected; the
// density-matrix Kraus map reaches it as a superoperator on 2k targets,
// distributed_densitymatrix.hpp:79-89, misc.hpp:58-81).
//
// The amplitudes of group g at target offset j form column g, row j of X
// (D = 2^k rows, 2^(n-k) columns); the gate is Y = U X.  A workgroup owns a
// column block of CBW groups: the groups whose non-target bits above the
// lowest log2(CBW) ones agree, so that a block is a union of contiguous runs.
// It stages the block in LDS in chunks of KC rows (4096 amplitudes per chunk,
// 16 per thread, all loads in flight at once), each in address order (a thread's low eight
// local-index bits are its id, so a wave reads consecutive amplitudes where
// the low bits are not targets).  Four waves split the work into units of 16
// columns x (16 RBW) rows.  The complex product takes three real MFMAs per
// 16-row block and k-step on v_mfma_*_16x16x4 (Gauss's three-multiplication
// form): P1 = Ur Xr, P2 = Ui Xi and P3 = (Ur + Ui)(Xr + Xi) accumulate
// apart, and Yr = P1 - P2, Yi = P3 - P1 - P2 at the end, a quarter fewer
// matrix operations than the four products of [[Ur, -Ui], [Ui, Ur]] [Xr; Xi]
// (FSA_GEMM_3M=0 keeps those).  The sums Ur + Ui and Xr + Xi are formed per
// lane as the operands arrive.  The A operand comes from the
// matrix packed in MFMA fragment order, one 2 x sizeof(R) load per lane and
// fragment from L2; the B operand is one LDS read of the staged amplitude.
//
// When one workgroup computes every row of its columns (D <= 256) it writes
// the result over its own staged block and stores it back in place.  Larger
// gates split the rows into panels over several workgroups, which all read
// the whole column block: their results go to a scratch area, scattered back
// into the state by k_gemm_scatter once every panel of the batch is done.
#include "launch.hpp"

#include <cstdlib>

namespace fsa {

namespace {

constexpr int GEMM_THREADS = 256;

template <typename R> struct GemmAcc;
template <> struct GemmAcc<double> {
    typedef double T __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ T mma(double a, double b, T c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator element j of lane l: row 4 j + l / 16, column l % 16
    static __device__ __forceinline__ uint32_t row(uint32_t j, uint32_t lk) { return 4u * j + lk; }
};
template <> struct GemmAcc<float> {
    typedef float T __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ T mma(float a, float b, T c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator element j of lane l: row 4 (l / 16) + j, column l % 16
    static __device__ __forceinline__ uint32_t row(uint32_t j, uint32_t lk) { return 4u * lk + j; }
};

// local index bits [LO, hi) of a staged element -> (address offset, LDS slot);
// unrolled so that the geometry is read with constant offsets (scalar loads)
template <int LO, int HI>
__device__ __forceinline__ void gemm_local(const GemmGeom& g, uint32_t x, uint32_t hi, uint64_t& addr, uint32_t& pos) {
    addr = 0;
    pos = 0;
#pragma unroll
    for (int b = LO; b < HI; b++)
        if ((uint32_t)b < hi && ((x >> (b - LO)) & 1u)) {
            addr |= 1ULL << g.lpos[b];
            pos += g.lw[b];
        }
}

template <typename R, int RBW, bool G3>
__global__ __launch_bounds__(GEMM_THREADS, 2) void k_matk_gemm(cx<R>* __restrict__ a,
                                                              const GemmFrag<R>* __restrict__ frag, GemmGeom g,
                                                              uint64_t cblkFirst, cx<R>* __restrict__ scratch) {
    typedef GemmAcc<R> A;
    typedef typename A::T acc_t;
    constexpr int NHI = (int)(GEMM_CHUNK / GEMM_THREADS);      // staged elements per thread
    __shared__ cx<R> xs[GEMM_CHUNK];
    __shared__ uint64_t hiAddr[NHI];
    __shared__ uint32_t hiPos[NHI];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t lr = lane & 15u, lk = lane >> 4;
    const uint32_t D = 1u << g.k, KC = 1u << g.kcBits, CBW = 1u << g.cbits;
    // a chunk holds GEMM_CHUNK elements: local index bits 0-7 from the thread id,
    // 8-11 from the element number h
    const uint32_t panel = blockIdx.x % g.panels;
    const uint64_t cblkLocal = blockIdx.x / g.panels;
    const uint64_t base = ins0_mask(cblkFirst + cblkLocal, g.fixmask);

    uint64_t loAddr;
    uint32_t loPos;
    gemm_local<0, 8>(g, tid, 8, loAddr, loPos);
    if (tid < (uint32_t)NHI) {
        uint64_t ad;
        uint32_t ps;
        gemm_local<8, 12>(g, tid, 12, ad, ps);
        hiAddr[tid] = ad;
        hiPos[tid] = ps;
    }

    // this wave's unit: RBW 16-row blocks of the panel, 16 columns
    const uint32_t ru = wave % g.rowUnits, cu = wave / g.rowUnits;
    const uint32_t rbFirst = (panel * g.rowUnits + ru) * RBW;
    const uint32_t colInBlk = cu * 16u + lr;
    // G3: yr, yi, ys accumulate P1, P2, P3; otherwise yr, yi are Yr, Yi
    acc_t yr[RBW], yi[RBW], ys[RBW];
#pragma unroll
    for (int r = 0; r < RBW; r++) yr[r] = yi[r] = ys[r] = acc_t{0, 0, 0, 0};
    const uint32_t kPerRb = D / 4;                          // fragments per 16-row block
    const uint32_t nChunks = D >> g.kcBits;
    const uint32_t nks = KC / 4;                            // k-steps per chunk: 4 or more, even

    for (uint32_t c = 0; c < nChunks; c++) {
        uint64_t chunkOff = 0;
#pragma unroll
        for (int i = 0; i < 12; i++)
            if ((uint32_t)i >= g.kcBits && (uint32_t)i < g.k && ((c >> ((uint32_t)i - g.kcBits)) & 1u))
                chunkOff |= 1ULL << g.tq[i];
        __syncthreads();          // the tables are written / the previous chunk is consumed
        // every load of the chunk in flight before the first LDS store
        const cx<R>* src = a + (base | chunkOff | loAddr);
        cx<R> v[NHI];
#pragma unroll
        for (int h = 0; h < NHI; h++) v[h] = src[hiAddr[h]];
#pragma unroll
        for (int h = 0; h < NHI; h++) xs[loPos + hiPos[h]] = v[h];
        __syncthreads();
        // two register sets of A fragments and B values: the loads of k-step
        // ks + 1 are in flight while the MFMAs of ks run
        const GemmFrag<R>* fr = frag + ((size_t)rbFirst * kPerRb + (size_t)c * nks) * 64 + lane;
        const cx<R>* xb = xs + lk * CBW + colInBlk;
        GemmFrag<R> m0[RBW], m1[RBW];
        cx<R> x0, x1;
        auto load = [&](GemmFrag<R>* m, cx<R>& x, uint32_t ks) {
#pragma unroll
            for (int r = 0; r < RBW; r++) m[r] = fr[((size_t)r * kPerRb + ks) * 64];
            x = xb[ks * 4u * CBW];
        };
        auto step = [&](const GemmFrag<R>* m, const cx<R>& x) {
            if constexpr (G3) {
                const R xsum = x.x + x.y;
#pragma unroll
                for (int r = 0; r < RBW; r++) {
                    yr[r] = A::mma(m[r].re, x.x, yr[r]);
                    yi[r] = A::mma(m[r].im, x.y, yi[r]);
                    ys[r] = A::mma(m[r].re + m[r].im, xsum, ys[r]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < RBW; r++) {
                    yr[r] = A::mma(m[r].re, x.x, yr[r]);
                    yi[r] = A::mma(m[r].im, x.x, yi[r]);
                    yr[r] = A::mma(-m[r].im, x.y, yr[r]);
                    yi[r] = A::mma(m[r].re, x.y, yi[r]);
                }
            }
        };
        load(m0, x0, 0);
#pragma unroll 1
        for (uint32_t ks = 0; ks < nks; ks += 2) {
            load(m1, x1, ks + 1);
            step(m0, x0);
            if (ks + 2 < nks) load(m0, x0, ks + 2);
            step(m1, x1);
        }
    }

    if constexpr (G3) {
#pragma unroll
        for (int r = 0; r < RBW; r++) {
            const acc_t p1 = yr[r], p2 = yi[r];
            yr[r] = p1 - p2;
            yi[r] = ys[r] - p1 - p2;
        }
    }
    if (g.inplace) {
        // one panel, one chunk: the block is this workgroup's alone
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RBW; r++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t row = (rbFirst + (uint32_t)r) * 16u + A::row(j, lk);
                xs[row * CBW + colInBlk] = mk<R>(yr[r][j], yi[r][j]);
            }
        __syncthreads();
        cx<R>* dst = a + (base | loAddr);
#pragma unroll
        for (int h = 0; h < NHI; h++) st_stream(dst + hiAddr[h], xs[loPos + hiPos[h]]);
    } else {
        cx<R>* out = scratch + cblkLocal * ((uint64_t)D * CBW);
#pragma unroll
        for (int r = 0; r < RBW; r++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t row = (rbFirst + (uint32_t)r) * 16u + A::row(j, lk);
                out[(uint64_t)row * CBW + colInBlk] = mk<R>(yr[r][j], yi[r][j]);
            }
    }
}

// scratch rows -> the state: element e of the batch is (column block e >> fb,
// block-local index over the full tile bits: every target and the column bits)
template <typename R>
__global__ __launch_bounds__(GEMM_THREADS) void k_gemm_scatter(cx<R>* __restrict__ a, const cx<R>* __restrict__ scratch,
                                                               GemmGeom g, uint64_t cblkFirst, uint64_t numElems) {
    const uint64_t e = (uint64_t)blockIdx.x * GEMM_THREADS + threadIdx.x;
    if (e >= numElems) return;
    const uint32_t fb = g.k + g.cbits;
    const uint64_t cblkLocal = e >> fb;
    const uint32_t loc = (uint32_t)(e & ((1ULL << fb) - 1));
    uint64_t addr = ins0_mask(cblkFirst + cblkLocal, g.fixmask);
    uint32_t pos = 0;
#pragma unroll
    for (int b = 0; b < 20; b++)
        if ((uint32_t)b < fb && ((loc >> b) & 1u)) {
            addr |= 1ULL << g.fpos[b];
            pos += g.fw[b];
        }
    a[addr] = scratch[(cblkLocal << fb) + pos];
}

}  // namespace

GemmGeom gemm_geom(unsigned nq, const unsigned* targets, unsigned k, size_t realBytes) {
    GemmGeom g{};
    const uint32_t D = 1u << k;
    g.k = k;
    g.rowUnits = D >= 256 ? 4u : D >= 128 ? 2u : 1u;
    const uint32_t rbw = D >= 64 ? 4u : D / 16u;            // 16-row blocks per wave
    const uint32_t cpw = 4u / g.rowUnits;                   // 16-column units per workgroup
    g.cbits = 4u + (cpw == 4 ? 2u : cpw == 2 ? 1u : 0u);
    g.kcBits = std::min<uint32_t>(k, 12u - g.cbits);        // KC * CBW = GEMM_CHUNK for k >= 6
    g.panels = D / (16u * rbw * g.rowUnits);
    g.rbw = rbw;
    uint64_t tmask = 0;
    for (unsigned i = 0; i < k; i++) {
        tmask |= 1ULL << targets[i];
        g.tq[i] = (uint8_t)targets[i];
    }
    // column bits: the lowest non-target qubits
    uint64_t cmask = 0;
    uint32_t cpos[8];
    for (unsigned q = 0, n = 0; q < nq && n < g.cbits; q++)
        if (!((tmask >> q) & 1)) { cmask |= 1ULL << q; cpos[n++] = q; }
    g.fixmask = tmask | cmask;
    g.inplace = (g.panels == 1 && g.kcBits == k) ? 1u : 0u;
    // chunk tile: the first kcBits targets and the column bits, ascending;
    // LDS slot = row * CBW + column
    auto weight = [&](unsigned q, unsigned rowBits) -> int64_t {
        for (unsigned i = 0; i < rowBits; i++) if (targets[i] == q) return 1LL << (g.cbits + i);
        for (unsigned c = 0; c < g.cbits; c++) if (cpos[c] == q) return 1LL << c;
        return -1;
    };
    unsigned n = 0;
    for (unsigned q = 0; q < nq; q++) {
        int64_t w = weight(q, g.kcBits);
        if (w < 0) continue;
        g.lpos[n] = (uint8_t)q;
        g.lw[n] = (uint32_t)w;
        n++;
    }
    n = 0;
    for (unsigned q = 0; q < nq; q++) {
        int64_t w = weight(q, k);
        if (w < 0) continue;
        g.fpos[n] = (uint8_t)q;
        g.fw[n] = (uint32_t)w;
        n++;
    }
    return g;
}

bool gemm_fits(unsigned nq, unsigned k) {
    if (k < 6 || k > 12) return false;
    const uint32_t D = 1u << k;
    const uint32_t rowUnits = D >= 256 ? 4u : D >= 128 ? 2u : 1u;
    const uint32_t cbits = 4u + (rowUnits == 1 ? 2u : rowUnits == 2 ? 1u : 0u);
    return nq >= k + cbits;
}

template <typename R>
void gemm_pack_matrix(const std::complex<double>* m, unsigned k, GemmFrag<R>* out) {
    const uint64_t D = 1ULL << k, kPerRb = D / 4;
    for (uint64_t rb = 0; rb < D / 16; rb++)
        for (uint64_t ks = 0; ks < kPerRb; ks++)
            for (uint32_t l = 0; l < 64; l++) {
                // A[l % 16][l / 16] of block (rb, ks)
                const std::complex<double> v = m[(rb * 16 + (l & 15)) * D + ks * 4 + (l >> 4)];
                GemmFrag<R>& f = out[(rb * kPerRb + ks) * 64 + l];
                f.re = (R)v.real();
                f.im = (R)v.imag();
            }
}

// FSA_GEMM_3M=0: four real products per complex one instead of three
static bool gemm_three_mult() {
    const char* e = std::getenv("FSA_GEMM_3M");
    return !(e && e[0] == '0');
}

uint64_t gemm_scratch_elems(const GemmGeom& g, unsigned nq, size_t realBytes) {
    if (g.inplace) return 0;
    const uint64_t perBlk = 1ULL << (g.k + g.cbits);
    const uint64_t numCblk = 1ULL << (nq - g.k - g.cbits);
    // FSA_GEMM_SCRATCH_KB caps the area (default 64 MiB); batches of column
    // blocks fill it
    const char* env = std::getenv("FSA_GEMM_SCRATCH_KB");
    const uint64_t capBytes = (env && *env ? std::strtoull(env, nullptr, 10) : 65536ULL) << 10;
    const uint64_t cap = std::max<uint64_t>(1, capBytes / (perBlk * 2 * realBytes));
    return std::min(numCblk, cap) * perBlk;
}

template <typename R>
void launch_matk_gemm(hipStream_t s, cx<R>* a, unsigned nq, const GemmGeom& g, const GemmFrag<R>* dfrag,
                      cx<R>* scratch) {
    const uint64_t numCblk = 1ULL << (nq - g.k - g.cbits);
    // k >= 6: four 16-row blocks per wave
    auto kern = gemm_three_mult() ? k_matk_gemm<R, 4, true> : k_matk_gemm<R, 4, false>;
    if (g.inplace) {
        hipLaunchKernelGGL(kern, dim3((unsigned)numCblk), dim3(GEMM_THREADS), 0, s, a, dfrag, g, (uint64_t)0,
                           (cx<R>*)nullptr);
        return;
    }
    const uint64_t perBlk = 1ULL << (g.k + g.cbits);
    const uint64_t batch = gemm_scratch_elems(g, nq, sizeof(R)) / perBlk;
    for (uint64_t first = 0; first < numCblk; first += batch) {
        const uint64_t nb = std::min(batch, numCblk - first);
        hipLaunchKernelGGL(kern, dim3((unsigned)(nb * g.panels)), dim3(GEMM_THREADS), 0, s, a, dfrag, g, first,
                           scratch);
        const uint64_t elems = nb * perBlk;
        hipLaunchKernelGGL(k_gemm_scatter<R>, dim3((unsigned)((elems + GEMM_THREADS - 1) / GEMM_THREADS)),
                           dim3(GEMM_THREADS), 0, s, a, (const cx<R>*)scratch, g, first, elems);
    }
}

template void gemm_pack_matrix<double>(const std::complex<double>*, unsigned, GemmFrag<double>*);
template void gemm_pack_matrix<float>(const std::complex<double>*, unsigned, GemmFrag<float>*);
template void launch_matk_gemm<double>(hipStream_t, cx<double>*, unsigned, const GemmGeom&, const GemmFrag<double>*,
                                       cx<double>*);
template void launch_matk_gemm<float>(hipStream_t, cx<float>*, unsigned, const GemmGeom&, const GemmFrag<float>*,
                                      cx<float>*);

}  // namespace fsa
