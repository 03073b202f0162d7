// noalloc.hip -- expecPauliString without any device or pinned allocation.
//
// fsa_densitymatrix_expecPauliString keeps its partial sums in the state's
// reduction workspace and its term list in the state's parameter ring, both
// grown with hipMalloc on demand.  The variant here takes that scratch from
// `scratch`, a file-scope device array of NOALLOC_SCRATCH_BYTES that every
// GPU holds (its own copy) until the process ends, so a caller that cannot
// allocate (code under HIP graph capture, a serving loop, a harness that
// forbids hipMalloc) can evaluate <H>.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (two host threads, two streams) race.
#include "engine.hpp"

#include <mutex>

namespace fsa {

namespace {

constexpr uint64_t NOALLOC_SCRATCH_BYTES = 4ULL << 20;
constexpr uint64_t NOALLOC_SCRATCH_DOUBLES = NOALLOC_SCRATCH_BYTES / sizeof(double);
constexpr int NOALLOC_MAX_DEVICES = 64;

// Host code never passes `scratch` itself to a kernel or copy: in host code the
// name is a stand-in, not a device address.  scratch_ptr() returns the address
// of the current GPU's copy.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

double* scratch_ptr() {
    static std::mutex mu;
    static double* ptrs[NOALLOC_MAX_DEVICES] = {};
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    need(dev >= 0 && dev < NOALLOC_MAX_DEVICES, FSA_ERR_SIZE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        need(p != nullptr, FSA_ERR_HIP, "no device address for the scratch array");
        ptrs[dev] = static_cast<double*>(p);
    }
    return ptrs[dev];
}

// term t from its coefficient and N Pauli codes (codes[t * N + q] on qubit q):
// masks, and f = coeff * (-i)^{#Y} (the density-matrix phase of
// distributed_densitymatrix.hpp:322-344, as pauli_terms builds it on the host)
__global__ void k_dm_terms(const double* __restrict__ coeffs, const uint32_t* __restrict__ codes, unsigned T,
                           unsigned N, PauliTermDev* __restrict__ out) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint64_t mxy = 0, myz = 0;
    unsigned ny = 0;
    for (unsigned q = 0; q < N; q++) {
        const uint32_t c = codes[(uint64_t)t * N + q];
        if (c == 1 || c == 2) mxy |= 1ULL << q;
        if (c == 2 || c == 3) myz |= 1ULL << q;
        if (c == 2) ny++;
    }
    // (-i)^ny: 1, -i, -1, i
    const double re[4] = {1.0, 0.0, -1.0, 0.0}, im[4] = {0.0, -1.0, 0.0, 1.0};
    out[t] = PauliTermDev{mxy, myz, coeffs[t] * re[ny & 3u], coeffs[t] * im[ny & 3u]};
}

size_t up32(size_t doubles) { return (doubles + 31) & ~(size_t)31; }

template <typename R>
void expec_noalloc(State& s, const double* coeffs, unsigned T, const unsigned* paulis, double* out) {
    double* scr = scratch_ptr();
    // layout (doubles, each region 256-byte aligned): coefficients, codes,
    // terms, partial sums, the result
    const size_t oCodes = up32(T);
    const size_t oTerms = oCodes + up32(((size_t)T * s.N + 1) / 2);
    const size_t oPart = oTerms + up32((size_t)T * (sizeof(PauliTermDev) / sizeof(double)));
    const size_t oRes = oPart + up32(expec_workspace_doubles(s.N, T));
    need(oRes + 2 <= NOALLOC_SCRATCH_DOUBLES, FSA_ERR_SIZE,
         "expecPauliString without allocation: terms x qubits exceed the 4 MiB scratch; use the allocating call");
    hip_check(hipMemcpyAsync(scr, coeffs, T * sizeof(double), hipMemcpyHostToDevice, s.stream), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(scr + oCodes, paulis, (size_t)T * s.N * sizeof(unsigned), hipMemcpyHostToDevice,
                             s.stream), "hipMemcpyAsync");
    auto* terms = reinterpret_cast<PauliTermDev*>(scr + oTerms);
    hipLaunchKernelGGL(k_dm_terms, dim3((T + 255) / 256), dim3(256), 0, s.stream, scr,
                       reinterpret_cast<const uint32_t*>(scr + oCodes), T, s.N, terms);
    launch_dm_expec<R>(s.stream, static_cast<const cx<R>*>(s.amps), s.N, terms, T, scr + oPart, scr + oRes);
    hip_check(hipGetLastError(), "expec kernel");
    hip_check(hipMemcpyAsync(s.hres, scr + oRes, 2 * sizeof(double), hipMemcpyDeviceToHost, s.stream),
              "hipMemcpyAsync(result)");
    hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    out[0] = s.hres[0];
    out[1] = s.hres[1];
}

}  // namespace
}  // namespace fsa

using namespace fsa;

extern "C" int fsa_densitymatrix_expecPauliString_noalloc(fsa_state* h, const double* coeffs, unsigned numTerms,
                                                          const unsigned* allPaulis, double* out) {
    return guard([&] {
        need(h != nullptr, FSA_ERR_ARG, "null state");
        State& s = h->s;
        need(s.dm, FSA_ERR_KIND, "density-matrix operation on a statevector");
        need(out != nullptr && (numTerms == 0 || (coeffs && allPaulis)), FSA_ERR_ARG, "null argument");
        for (uint64_t i = 0; i < (uint64_t)numTerms * s.N; i++)
            need(allPaulis[i] <= 3, FSA_ERR_ARG, "Pauli codes are 0..3 (I, X, Y, Z)");
        // applying queued operations may grow the parameter ring (hipMalloc)
        need(s.pending.empty(), FSA_ERR_ARG,
             "the state has queued operations: flush it before the allocation-free call");
        hipPointerAttribute_t attr;
        hip_check(hipPointerGetAttributes(&attr, s.amps), "hipPointerGetAttributes");
        int dev = -1;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        need(attr.device == dev, FSA_ERR_ARG, "the state's device must be the current device");
        if (numTerms == 0) { out[0] = out[1] = 0; return; }
        if (s.prec == FSA_FP64) expec_noalloc<double>(s, coeffs, numTerms, allPaulis, out);
        else expec_noalloc<float>(s, coeffs, numTerms, allPaulis, out);
    });
}
