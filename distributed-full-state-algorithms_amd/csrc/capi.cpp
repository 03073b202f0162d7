// capi.cpp -- the extern "C" boundary of libfsa (declared in include/fsa.h).
//
// Each entry point validates its arguments (the reference asserts; this
// library returns a status and keeps the message for fsa_last_error), turns
// the call into primitive operations on the amplitude vector, and either runs
// them (eager mode) or queues them for the next fused flush.
#include "../../include/fsa.h"
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace fsa;

namespace {

thread_local std::string g_err;
bool g_depol2_as_shipped = false;

State& st(fsa_state* s) {
    need(s != nullptr, FSA_ERR_ARG, "null state");
    return s->s;
}
State& sv(fsa_state* s) {
    State& x = st(s);
    need(!x.dm, FSA_ERR_KIND, "statevector operation on a density matrix");
    return x;
}
State& dmx(fsa_state* s) {
    State& x = st(s);
    need(x.dm, FSA_ERR_KIND, "density-matrix operation on a statevector");
    return x;
}

void check_qubits(const unsigned* q, unsigned num, unsigned N, const char* what) {
    need(num == 0 || q != nullptr, FSA_ERR_ARG, std::string("null ") + what);
    uint64_t seen = 0;
    for (unsigned i = 0; i < num; i++) {
        need(q[i] < N, FSA_ERR_QUBIT, std::string(what) + " index out of range");
        need(!((seen >> q[i]) & 1), FSA_ERR_DUPLICATE, std::string("repeated qubit among ") + what);
        seen |= 1ULL << q[i];
    }
}

std::vector<cd> read_mat(const double* g, uint64_t dim) {
    need(g != nullptr, FSA_ERR_ARG, "null gate matrix");
    std::vector<cd> m(dim * dim);
    for (uint64_t i = 0; i < dim * dim; i++) m[i] = cd(g[2 * i], g[2 * i + 1]);
    return m;
}

void check_prob(double p) { need(std::isfinite(p), FSA_ERR_ARG, "probability must be finite"); }

// ---- primitive op builders --------------------------------------------------

HOp op_mat1(unsigned t, uint64_t ctrl, std::vector<cd> m) {
    HOp o;
    o.kind = OP_MAT1;
    o.q = {t};
    o.ctrl = ctrl;
    o.mat = std::move(m);
    finish_op(o);
    return o;
}

HOp op_matk(std::vector<unsigned> targets, std::vector<cd> m) {
    if (targets.size() == 1) return op_mat1(targets[0], 0, std::move(m));
    HOp o;
    o.kind = OP_MATK;
    o.q = std::move(targets);
    o.mat = std::move(m);
    finish_op(o);
    return o;
}

HOp op_swap(unsigned a, unsigned b) {
    HOp o;
    o.kind = OP_SWAP;
    o.q = {std::min(a, b), std::max(a, b)};
    finish_op(o);
    return o;
}

// reference distributed_statevector.hpp:244-276: powI, maskXY, maskYZ
HOp op_pauli(const unsigned* targets, const unsigned* paulis, unsigned num, unsigned shift, bool gadget,
             double theta, double otherSign) {
    HOp o;
    o.kind = OP_PAULI;
    o.pw = 1.0;
    for (unsigned i = 0; i < num; i++) {
        unsigned q = targets[i] + shift;
        if (paulis[i] == 2) o.pw *= cd(0, 1);
        if (paulis[i] == 1 || paulis[i] == 2) o.mxy |= 1ULL << q;
        if (paulis[i] == 2 || paulis[i] == 3) o.myz |= 1ULL << q;
    }
    o.gadget = gadget;
    o.theta = theta;
    o.other_sign = otherSign;
    refresh_angle(o);
    finish_op(o);
    return o;
}

HOp op_phase(const unsigned* targets, unsigned num, unsigned shift, double theta) {
    HOp o;
    o.kind = OP_PHASE;
    for (unsigned i = 0; i < num; i++) o.pmask |= 1ULL << (targets[i] + shift);
    o.theta = theta;
    finish_op(o);
    return o;
}

HOp op_chan(OpKind k, std::vector<unsigned> q, double c0, double c1, double c2) {
    HOp o;
    o.kind = k;
    o.q = std::move(q);
    o.c[0] = c0; o.c[1] = c1; o.c[2] = c2;
    finish_op(o);
    return o;
}

void check_paulis(const unsigned* p, unsigned num) {
    need(num == 0 || p != nullptr, FSA_ERR_ARG, "null pauli codes");
    for (unsigned i = 0; i < num; i++) need(p[i] <= 3, FSA_ERR_ARG, "Pauli codes are 0..3 (I, X, Y, Z)");
}

bool odd_y(const unsigned* p, unsigned num) {
    bool odd = false;
    for (unsigned i = 0; i < num; i++) if (p[i] == 2) odd = !odd;   // misc.hpp:47-55
    return odd;
}

void submit(State& s, std::vector<HOp> ops) {
    if (s.mode == FSA_EXEC_EAGER) {
        execute_now(s, ops);
        return;
    }
    for (HOp& o : ops) s.pending.push_back(std::move(o));
    if (s.pending.size() >= 4096) flush(s);
}

void readback(State& s, double* dres, double* out, int n) {
    hip_check(hipMemcpyAsync(s.hres, dres, n * sizeof(double), hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync(result)");
    hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    for (int i = 0; i < n; i++) out[i] = s.hres[i];
}

std::vector<PauliTermDev> pauli_terms(const double* coeffs, unsigned T, const unsigned* paulis, unsigned n,
                                      bool dm) {
    std::vector<PauliTermDev> terms(T);
    for (unsigned t = 0; t < T; t++) {
        uint64_t mxy = 0, myz = 0;
        cd ph = 1.0;
        for (unsigned q = 0; q < n; q++) {
            unsigned c = paulis[(uint64_t)t * n + q];
            if (c == 1 || c == 2) mxy |= 1ULL << q;
            if (c == 2 || c == 3) myz |= 1ULL << q;
            if (c == 2) ph *= dm ? cd(0, -1) : cd(0, 1);
        }
        cd f = coeffs[t] * ph;
        terms[t] = PauliTermDev{mxy, myz, f.real(), f.imag()};
    }
    return terms;
}

void create_state(unsigned N, bool dm, int prec, fsa_state** out) {
    need(out != nullptr, FSA_ERR_ARG, "null output handle");
    need(prec == FSA_FP64 || prec == FSA_FP32, FSA_ERR_ARG, "precision must be FSA_FP64 or FSA_FP32");
    unsigned nq = dm ? 2 * N : N;
    need(nq <= 40, FSA_ERR_SIZE, "too many qubits for one device");
    auto* h = new fsa_state();
    State& s = h->s;
    s.N = N;
    s.nq = nq;
    s.dm = dm;
    s.prec = prec;
    try {
        hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
        s.ring.stream = s.stream;
        hip_check(hipMalloc(&s.amps, s.num() * s.amp_bytes()), "hipMalloc(amplitudes)");
        hip_check(hipHostMalloc((void**)&s.hres, 64 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        if (prec == FSA_FP64) launch_init_basis<double>(s.stream, (cx<double>*)s.amps, s.num(), s.num());
        else launch_init_basis<float>(s.stream, (cx<float>*)s.amps, s.num(), s.num());
        hip_check(hipGetLastError(), "zero fill");
    } catch (...) {
        fsa_state_destroy(h);
        throw;
    }
    *out = h;
}

constexpr uint64_t CONV_CHUNK = 1ULL << 22;

}  // namespace

void fsa::set_last_error(const std::string& msg) { g_err = msg; }

// =============================================================================

extern "C" {

const char* fsa_last_error(void) { return g_err.c_str(); }
int fsa_version(void) { return 1; }

int fsa_comm_init(void) {
    return guard([] {
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        need(n > 0, FSA_ERR_HIP, "no HIP device");
    });
}
int fsa_comm_end(void) { return guard([] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }); }
unsigned fsa_comm_getRank(void) { return 0; }
unsigned fsa_comm_getNumNodes(void) { return 1; }
int fsa_comm_synch(void) { return guard([] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }); }
int fsa_comm_reduceAmp(double* amp) { return guard([&] { need(amp != nullptr, FSA_ERR_ARG, "null amp"); }); }

int fsa_profile_enable(int enable) {
    return guard([&] { profile_enable(enable != 0); });
}

int fsa_profile_read(fsa_profile_entry* out, unsigned capacity, unsigned* numEntries) {
    return guard([&] {
        need(numEntries != nullptr, FSA_ERR_ARG, "null output");
        auto entries = profile_read();
        *numEntries = (unsigned)entries.size();
        for (unsigned i = 0; i < entries.size() && i < capacity && out; i++) {
            std::memset(out[i].name, 0, sizeof(out[i].name));
            std::strncpy(out[i].name, entries[i].name.c_str(), sizeof(out[i].name) - 1);
            out[i].count = entries[i].count;
            out[i].total_ms = entries[i].ms;
        }
    });
}

int fsa_set_two_qubit_depolarising_as_shipped(int enable) {
    g_depol2_as_shipped = enable != 0;
    return FSA_OK;
}

// ---- states -------------------------------------------------------------------

int fsa_state_create(unsigned numQubits, int isDensityMatrix, int precision, fsa_state** out) {
    return guard([&] { create_state(numQubits, isDensityMatrix != 0, precision, out); });
}

int fsa_state_destroy(fsa_state* h) {
    if (!h) return FSA_OK;
    State& s = h->s;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.amps) (void)hipFree(s.amps);
    if (s.ws) (void)hipFree(s.ws);
    if (s.hres) (void)hipHostFree(s.hres);
    // the ring frees its buffers in its destructor; the stream goes last
    hipStream_t stream = s.stream;
    delete h;
    if (stream) (void)hipStreamDestroy(stream);
    return FSA_OK;
}

int fsa_state_info(const fsa_state* h, unsigned* numQubits, int* isDm, int* precision, uint64_t* numAmps) {
    return guard([&] {
        need(h != nullptr, FSA_ERR_ARG, "null state");
        if (numQubits) *numQubits = h->s.N;
        if (isDm) *isDm = h->s.dm ? 1 : 0;
        if (precision) *precision = h->s.prec;
        if (numAmps) *numAmps = h->s.num();
    });
}

int fsa_state_set_exec_mode(fsa_state* h, int mode) {
    return guard([&] {
        State& s = st(h);
        need(mode == FSA_EXEC_EAGER || mode == FSA_EXEC_FUSED, FSA_ERR_ARG, "unknown execution mode");
        flush(s);
        s.mode = mode;
    });
}

int fsa_state_flush(fsa_state* h) { return guard([&] { flush(st(h)); }); }

int fsa_state_sync(fsa_state* h) {
    return guard([&] {
        State& s = st(h);
        flush(s);
        hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    });
}

int fsa_state_upload(fsa_state* h, const double* amps, uint64_t numAmps) {
    return guard([&] {
        State& s = st(h);
        need(amps != nullptr, FSA_ERR_ARG, "null amplitudes");
        need(numAmps == s.num(), FSA_ERR_ARG, "amplitude count does not match the state");
        s.pending.clear();   // overwritten anyway
        if (s.prec == FSA_FP64) {
            hip_check(hipMemcpyAsync(s.amps, amps, s.num() * 16, hipMemcpyHostToDevice, s.stream), "hipMemcpyAsync(upload)");
            hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            return;
        }
        cx<double>* tmp = nullptr;
        uint64_t chunk = std::min<uint64_t>(CONV_CHUNK, s.num());
        hip_check(hipMalloc((void**)&tmp, chunk * 16), "hipMalloc(conversion)");
        try {
            for (uint64_t o = 0; o < s.num(); o += chunk) {
                hip_check(hipMemcpyAsync(tmp, amps + 2 * o, chunk * 16, hipMemcpyHostToDevice, s.stream), "hipMemcpyAsync");
                launch_f64_to_f32(s.stream, tmp, (cx<float>*)s.amps + o, chunk);
                hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            }
        } catch (...) {
            (void)hipFree(tmp);
            throw;
        }
        hip_check(hipFree(tmp), "hipFree");
    });
}

int fsa_state_download(fsa_state* h, double* amps, uint64_t numAmps) {
    return guard([&] {
        State& s = st(h);
        need(amps != nullptr, FSA_ERR_ARG, "null amplitudes");
        need(numAmps == s.num(), FSA_ERR_ARG, "amplitude count does not match the state");
        flush(s);
        if (s.prec == FSA_FP64) {
            hip_check(hipMemcpyAsync(amps, s.amps, s.num() * 16, hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync(download)");
            hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            return;
        }
        cx<double>* tmp = nullptr;
        uint64_t chunk = std::min<uint64_t>(CONV_CHUNK, s.num());
        hip_check(hipMalloc((void**)&tmp, chunk * 16), "hipMalloc(conversion)");
        try {
            for (uint64_t o = 0; o < s.num(); o += chunk) {
                launch_f32_to_f64(s.stream, (const cx<float>*)s.amps + o, tmp, chunk);
                hip_check(hipMemcpyAsync(amps + 2 * o, tmp, chunk * 16, hipMemcpyDeviceToHost, s.stream), "hipMemcpyAsync");
                hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
            }
        } catch (...) {
            (void)hipFree(tmp);
            throw;
        }
        hip_check(hipFree(tmp), "hipFree");
    });
}

int fsa_state_copy(fsa_state* dst, fsa_state* src) {
    return guard([&] {
        State& d = st(dst);
        State& s = st(src);
        need(d.nq == s.nq && d.dm == s.dm && d.prec == s.prec, FSA_ERR_ARG, "states differ in shape or precision");
        flush(s);
        d.pending.clear();
        hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
        hip_check(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
        hip_check(hipMemcpyAsync(d.amps, s.amps, s.num() * s.amp_bytes(), hipMemcpyDeviceToDevice, d.stream), "hipMemcpyAsync(copy)");
        // later work on src (its own stream) must not overwrite the source
        // while the copy still reads it
        hipEvent_t done = nullptr;
        hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreate");
        hipError_t e = hipEventRecord(done, d.stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(s.stream, done, 0);
        (void)hipEventDestroy(done);
        hip_check(e, "ordering the copy before later work on the source");
    });
}

int fsa_state_init_zero(fsa_state* h) {
    return guard([&] {
        State& s = st(h);
        s.pending.clear();
        if (s.prec == FSA_FP64) launch_init_basis<double>(s.stream, (cx<double>*)s.amps, s.num(), 0);
        else launch_init_basis<float>(s.stream, (cx<float>*)s.amps, s.num(), 0);
        hip_check(hipGetLastError(), "init");
    });
}

int fsa_state_init_gaussian(fsa_state* h, uint64_t seed) {
    return guard([&] {
        State& s = st(h);
        s.pending.clear();
        double* ws = s.workspace(reduce_workspace_doubles(s.num()) + 8);
        if (s.prec == FSA_FP64) {
            auto* a = (cx<double>*)s.amps;
            launch_init_gaussian<double>(s.stream, a, s.nq, seed);
            launch_norm2<double>(s.stream, a, s.num(), ws + 8, ws);
            launch_scale_by_inv_sqrt<double>(s.stream, a, s.num(), ws);
        } else {
            auto* a = (cx<float>*)s.amps;
            launch_init_gaussian<float>(s.stream, a, s.nq, seed);
            launch_norm2<float>(s.stream, a, s.num(), ws + 8, ws);
            launch_scale_by_inv_sqrt<float>(s.stream, a, s.num(), ws);
        }
        hip_check(hipGetLastError(), "init");
    });
}

int fsa_state_device_ptr(fsa_state* h, void** ptr) {
    return guard([&] {
        State& s = st(h);
        need(ptr != nullptr, FSA_ERR_ARG, "null output");
        flush(s);
        *ptr = s.amps;
    });
}

int fsa_state_stream(fsa_state* h, void** stream) {
    return guard([&] {
        State& s = st(h);
        need(stream != nullptr, FSA_ERR_ARG, "null output");
        *stream = (void*)s.stream;
    });
}

int fsa_state_norm2(fsa_state* h, double* out) {
    return guard([&] {
        State& s = st(h);
        need(out != nullptr, FSA_ERR_ARG, "null output");
        flush(s);
        double* ws = s.workspace(reduce_workspace_doubles(s.num()) + 8);
        if (s.prec == FSA_FP64) launch_norm2<double>(s.stream, (const cx<double>*)s.amps, s.num(), ws + 8, ws);
        else launch_norm2<float>(s.stream, (const cx<float>*)s.amps, s.num(), ws + 8, ws);
        double r[2];
        readback(s, ws, r, 2);
        *out = r[0];
    });
}

int fsa_statevector_renormalise(fsa_state* h) {
    return guard([&] {
        State& s = sv(h);
        flush(s);
        // <psi|psi> summed in FP64 on the device, then every amplitude scaled by
        // its inverse square root, both on the state's stream (no host round trip)
        double* ws = s.workspace(reduce_workspace_doubles(s.num()) + 8);
        if (s.prec == FSA_FP64) {
            launch_norm2<double>(s.stream, (const cx<double>*)s.amps, s.num(), ws + 8, ws);
            launch_scale_by_inv_sqrt<double>(s.stream, (cx<double>*)s.amps, s.num(), ws);
        } else {
            launch_norm2<float>(s.stream, (const cx<float>*)s.amps, s.num(), ws + 8, ws);
            launch_scale_by_inv_sqrt<float>(s.stream, (cx<float>*)s.amps, s.num(), ws);
        }
        hip_check(hipGetLastError(), "renormalise");
    });
}

// ---- statevector ops -------------------------------------------------------------

int fsa_statevector_oneTargGate(fsa_state* h, unsigned target, const double* gate) {
    return guard([&] {
        State& s = sv(h);
        check_qubits(&target, 1, s.N, "target");
        submit(s, {op_mat1(target, 0, read_mat(gate, 2))});
    });
}

int fsa_statevector_manyCtrlOneTargGate(fsa_state* h, const unsigned* controls, unsigned numControls,
                                        unsigned target, const double* gate) {
    return guard([&] {
        State& s = sv(h);
        std::vector<unsigned> all(controls, controls + (controls ? numControls : 0));
        need(all.size() == numControls, FSA_ERR_ARG, "null controls");
        all.push_back(target);
        check_qubits(all.data(), (unsigned)all.size(), s.N, "control/target");
        uint64_t cm = 0;
        for (unsigned i = 0; i < numControls; i++) cm |= 1ULL << controls[i];
        submit(s, {op_mat1(target, cm, read_mat(gate, 2))});
    });
}

int fsa_statevector_swapGate(fsa_state* h, unsigned qb1, unsigned qb2) {
    return guard([&] {
        State& s = sv(h);
        need(qb1 < s.N && qb2 < s.N, FSA_ERR_QUBIT, "qubit index out of range");
        if (qb1 == qb2) return;
        submit(s, {op_swap(qb1, qb2)});
    });
}

int fsa_statevector_manyTargGate(fsa_state* h, const unsigned* targets, unsigned numTargets, const double* gate) {
    return guard([&] {
        State& s = sv(h);
        need(numTargets >= 1, FSA_ERR_ARG, "manyTargGate needs at least one target");
        check_qubits(targets, numTargets, s.N, "target");
        need(numTargets <= 12, FSA_ERR_SIZE, "manyTargGate supports at most 12 targets");
        submit(s, {op_matk(std::vector<unsigned>(targets, targets + numTargets), read_mat(gate, 1ULL << numTargets))});
    });
}

int fsa_statevector_pauliTensor(fsa_state* h, const unsigned* targets, const unsigned* paulis, unsigned num) {
    return guard([&] {
        State& s = sv(h);
        check_qubits(targets, num, s.N, "target");
        check_paulis(paulis, num);
        submit(s, {op_pauli(targets, paulis, num, 0, false, 0.0, 1.0)});
    });
}

int fsa_statevector_pauliGadget(fsa_state* h, const unsigned* targets, const unsigned* paulis, unsigned num,
                                double theta) {
    return guard([&] {
        State& s = sv(h);
        check_qubits(targets, num, s.N, "target");
        check_paulis(paulis, num);
        submit(s, {op_pauli(targets, paulis, num, 0, true, theta, 1.0)});
    });
}

int fsa_statevector_phaseGadget(fsa_state* h, const unsigned* targets, unsigned num, double theta) {
    return guard([&] {
        State& s = sv(h);
        check_qubits(targets, num, s.N, "target");
        submit(s, {op_phase(targets, num, 0, theta)});
    });
}

// ---- density-matrix ops (U on rows, conj(U) on columns, distributed_densitymatrix.hpp) ----

int fsa_densitymatrix_manyTargGate(fsa_state* h, const unsigned* targets, unsigned numTargets, const double* gate) {
    return guard([&] {
        State& s = dmx(h);
        need(numTargets >= 1, FSA_ERR_ARG, "manyTargGate needs at least one target");
        check_qubits(targets, numTargets, s.N, "target");
        need(numTargets <= 12, FSA_ERR_SIZE, "manyTargGate supports at most 12 targets");
        std::vector<cd> m = read_mat(gate, 1ULL << numTargets);
        std::vector<cd> mc(m.size());
        for (size_t i = 0; i < m.size(); i++) mc[i] = std::conj(m[i]);
        std::vector<unsigned> t(targets, targets + numTargets), t2(t);
        for (unsigned& x : t2) x += s.N;
        submit(s, {op_matk(t, std::move(m)), op_matk(t2, std::move(mc))});
    });
}

int fsa_densitymatrix_swapGate(fsa_state* h, unsigned qb1, unsigned qb2) {
    return guard([&] {
        State& s = dmx(h);
        need(qb1 < s.N && qb2 < s.N, FSA_ERR_QUBIT, "qubit index out of range");
        if (qb1 == qb2) return;
        submit(s, {op_swap(qb1, qb2), op_swap(qb1 + s.N, qb2 + s.N)});
    });
}

int fsa_densitymatrix_pauliTensor(fsa_state* h, const unsigned* targets, const unsigned* paulis, unsigned num) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(targets, num, s.N, "target");
        check_paulis(paulis, num);
        // the odd-Y negation of distributed_densitymatrix.hpp:45-49 is folded into the column op
        double sign = odd_y(paulis, num) ? -1.0 : 1.0;
        submit(s, {op_pauli(targets, paulis, num, 0, false, 0.0, 1.0),
                   op_pauli(targets, paulis, num, s.N, false, 0.0, sign)});
    });
}

int fsa_densitymatrix_pauliGadget(fsa_state* h, const unsigned* targets, const unsigned* paulis, unsigned num,
                                  double theta) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(targets, num, s.N, "target");
        check_paulis(paulis, num);
        double th2 = odd_y(paulis, num) ? theta : -theta;     // :60-61
        submit(s, {op_pauli(targets, paulis, num, 0, true, theta, 1.0),
                   op_pauli(targets, paulis, num, s.N, true, th2, 1.0)});
    });
}

int fsa_densitymatrix_phaseGadget(fsa_state* h, const unsigned* targets, unsigned num, double theta) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(targets, num, s.N, "target");
        submit(s, {op_phase(targets, num, 0, theta), op_phase(targets, num, s.N, -theta)});
    });
}

int fsa_densitymatrix_krausMap(fsa_state* h, const double* krausOps, unsigned numOps, const unsigned* targets,
                               unsigned numTargets) {
    return guard([&] {
        State& s = dmx(h);
        need(numTargets >= 1 && numOps >= 1, FSA_ERR_ARG, "krausMap needs targets and operators");
        check_qubits(targets, numTargets, s.N, "target");
        need(numTargets <= 6, FSA_ERR_SIZE, "krausMap supports at most 6 targets");
        need(krausOps != nullptr, FSA_ERR_ARG, "null Kraus operators");
        // misc.hpp:58-81: superOp[i*d+k][j*d+l] += conj(K[i][j]) * K[k][l]
        const uint64_t d = 1ULL << numTargets, D = d * d;
        std::vector<cd> sup(D * D, 0.0);
        for (unsigned o = 0; o < numOps; o++) {
            const double* K = krausOps + 2 * o * d * d;
            auto k_at = [&](uint64_t r, uint64_t c) { return cd(K[2 * (r * d + c)], K[2 * (r * d + c) + 1]); };
            for (uint64_t i = 0; i < d; i++)
                for (uint64_t j = 0; j < d; j++) {
                    cd cij = std::conj(k_at(i, j));
                    for (uint64_t k = 0; k < d; k++)
                        for (uint64_t l = 0; l < d; l++) sup[(i * d + k) * D + (j * d + l)] += cij * k_at(k, l);
                }
        }
        std::vector<unsigned> ext(targets, targets + numTargets);
        for (unsigned i = 0; i < numTargets; i++) ext.push_back(targets[i] + s.N);
        submit(s, {op_matk(ext, std::move(sup))});
    });
}

int fsa_densitymatrix_oneQubitDephasing(fsa_state* h, unsigned qb, double prob) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(&qb, 1, s.N, "qubit");
        check_prob(prob);
        submit(s, {op_chan(OP_DEPHASE1, {qb, qb + s.N}, 1 - 2 * prob, 0, 0)});
    });
}

int fsa_densitymatrix_twoQubitDephasing(fsa_state* h, unsigned qb1, unsigned qb2, double prob) {
    return guard([&] {
        State& s = dmx(h);
        unsigned q[2] = {qb1, qb2};
        check_qubits(q, 2, s.N, "qubit");
        check_prob(prob);
        submit(s, {op_chan(OP_DEPHASE2, {qb1, qb1 + s.N, qb2, qb2 + s.N}, -4 * prob / 3, 0, 0)});
    });
}

int fsa_densitymatrix_oneQubitDepolarising(fsa_state* h, unsigned qb, double prob) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(&qb, 1, s.N, "qubit");
        check_prob(prob);
        // distributed_densitymatrix.hpp:106-108
        submit(s, {op_chan(OP_DEPOL1, {qb, qb + s.N}, 2 * prob / 3, 1 - 2 * prob / 3, 1 - 4 * prob / 3)});
    });
}

int fsa_densitymatrix_twoQubitDepolarising(fsa_state* h, unsigned qb1, unsigned qb2, double prob) {
    return guard([&] {
        State& s = dmx(h);
        unsigned q[2] = {qb1, qb2};
        check_qubits(q, 2, s.N, "qubit");
        check_prob(prob);
        if (qb1 > qb2) std::swap(qb1, qb2);     // :244-245
        // :247-249 constants; the block-diagonal coefficient is the textbook
        // 1-16p/15 unless the shipped 1-4p/5 was requested
        double c1 = g_depol2_as_shipped ? 1 - 4 * prob / 5 : 1 - 16 * prob / 15;
        submit(s, {op_chan(OP_DEPOL2, {qb1, qb2, qb1 + s.N, qb2 + s.N}, c1, 4 * prob / 15, -16 * prob / 15)});
    });
}

int fsa_densitymatrix_damping(fsa_state* h, unsigned qb, double prob) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(&qb, 1, s.N, "qubit");
        check_prob(prob);
        need(prob <= 1.0, FSA_ERR_ARG, "damping probability must be <= 1");
        submit(s, {op_chan(OP_DAMP, {qb, qb + s.N}, prob, std::sqrt(1 - prob), 1 - prob)});
    });
}

int fsa_densitymatrix_expecPauliString(fsa_state* h, const double* coeffs, unsigned numTerms,
                                       const unsigned* allPaulis, double* out) {
    return guard([&] {
        State& s = dmx(h);
        need(out != nullptr && (numTerms == 0 || (coeffs && allPaulis)), FSA_ERR_ARG, "null argument");
        check_paulis(allPaulis, numTerms * s.N);
        flush(s);
        if (numTerms == 0) { out[0] = out[1] = 0; return; }
        auto terms = pauli_terms(coeffs, numTerms, allPaulis, s.N, true);
        auto* dt = (const PauliTermDev*)s.ring.push(terms.data(), terms.size() * sizeof(PauliTermDev));
        double* ws = s.workspace(expec_workspace_doubles(s.N, numTerms) + 8);
        if (s.prec == FSA_FP64) launch_dm_expec<double>(s.stream, (const cx<double>*)s.amps, s.N, dt, numTerms, ws + 8, ws);
        else launch_dm_expec<float>(s.stream, (const cx<float>*)s.amps, s.N, dt, numTerms, ws + 8, ws);
        hip_check(hipGetLastError(), "expec kernel");
        readback(s, ws, out, 2);
    });
}

int fsa_densitymatrix_partialTrace(fsa_state* h, const unsigned* targets, unsigned numTargets, fsa_state** out) {
    return guard([&] {
        State& s = dmx(h);
        need(out != nullptr, FSA_ERR_ARG, "null output handle");
        check_qubits(targets, numTargets, s.N, "target");
        need(numTargets <= s.N, FSA_ERR_ARG, "cannot trace out more qubits than the matrix has");
        std::vector<unsigned> t(targets, targets + numTargets);
        std::sort(t.begin(), t.end());                      // :362
        flush(s);
        fsa_state* o = nullptr;
        create_state(s.N - numTargets, true, s.prec, &o);
        try {
            // the output's zero fill runs on its own stream: order it before the trace
            hip_check(hipStreamSynchronize(o->s.stream), "hipStreamSynchronize");
            if (s.prec == FSA_FP64)
                launch_partial_trace<double>(s.stream, (const cx<double>*)s.amps, (cx<double>*)o->s.amps, s.N, t.data(), numTargets);
            else
                launch_partial_trace<float>(s.stream, (const cx<float>*)s.amps, (cx<float>*)o->s.amps, s.N, t.data(), numTargets);
            hip_check(hipGetLastError(), "partial trace kernel");
            hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
        } catch (...) {
            fsa_state_destroy(o);
            throw;
        }
        *out = o;
    });
}

// ---- the reference's local subroutines ------------------------------------------------

int fsa_local_statevector_pauliTensorOrGadget(fsa_state* h, uint64_t maskXY, uint64_t maskYZ, const double* powI,
                                              const double* thisAmpFac, const double* otherAmpFac) {
    return guard([&] {
        State& s = sv(h);
        need(powI && thisAmpFac && otherAmpFac, FSA_ERR_ARG, "null factor");
        const uint64_t inRange = s.N >= 64 ? ~0ULL : (1ULL << s.N) - 1;
        need(!(maskXY & ~inRange) && !(maskYZ & ~inRange), FSA_ERR_QUBIT, "Pauli mask names a qubit beyond the state");
        // local_statevector.hpp:126-131: new[j0] = this*a0 + other*powI*(-1)^{par(j1 & maskYZ)}*a1
        HOp o;
        o.kind = OP_PAULI;
        o.mxy = maskXY;
        o.myz = maskYZ;
        o.pw = cd(powI[0], powI[1]);
        o.th = cd(thisAmpFac[0], thisAmpFac[1]);
        o.fe = cd(otherAmpFac[0], otherAmpFac[1]) * o.pw;
        finish_op(o);
        submit(s, {std::move(o)});
    });
}

int fsa_local_densitymatrix_oneQubitDepolarising(fsa_state* h, unsigned qb, double c1, double c2, double c3) {
    return guard([&] {
        State& s = dmx(h);
        check_qubits(&qb, 1, s.N, "qubit");
        need(std::isfinite(c1) && std::isfinite(c2) && std::isfinite(c3), FSA_ERR_ARG, "coefficients must be finite");
        submit(s, {op_chan(OP_DEPOL1, {qb, qb + s.N}, c1, c2, c3)});
    });
}

int fsa_local_densitymatrix_twoQubitDepolarising(fsa_state* h, unsigned q0, unsigned q1, unsigned q2, unsigned q3,
                                                 double c1, double c2, double c3) {
    return guard([&] {
        State& s = dmx(h);
        const unsigned q[4] = {q0, q1, q2, q3};
        check_qubits(q, 4, s.nq, "qubit");
        need(std::isfinite(c1) && std::isfinite(c2) && std::isfinite(c3), FSA_ERR_ARG, "coefficients must be finite");
        submit(s, {op_chan(OP_DEPOL2, {q0, q1, q2, q3}, c1, c2, c3)});
    });
}

// ---- extensions -----------------------------------------------------------------------

int fsa_statevector_expecPauliSum(fsa_state* h, const double* coeffs, unsigned numTerms, const unsigned* paulis,
                                  double* out) {
    return guard([&] {
        State& s = sv(h);
        need(out != nullptr && (numTerms == 0 || (coeffs && paulis)), FSA_ERR_ARG, "null argument");
        check_paulis(paulis, numTerms * s.N);
        flush(s);
        if (numTerms == 0) { out[0] = out[1] = 0; return; }
        auto terms = pauli_terms(coeffs, numTerms, paulis, s.N, false);
        auto* dt = (const PauliTermDev*)s.ring.push(terms.data(), terms.size() * sizeof(PauliTermDev));
        double* ws = s.workspace(expec_workspace_doubles(s.N, numTerms) + 8);
        if (s.prec == FSA_FP64) launch_sv_expec<double>(s.stream, (const cx<double>*)s.amps, s.N, dt, numTerms, ws + 8, ws);
        else launch_sv_expec<float>(s.stream, (const cx<float>*)s.amps, s.N, dt, numTerms, ws + 8, ws);
        hip_check(hipGetLastError(), "expec kernel");
        readback(s, ws, out, 2);
    });
}

int fsa_statevector_applyPauliSum(fsa_state* h, const double* coeffs, unsigned numTerms, const unsigned* paulis,
                                  fsa_state* outState) {
    return guard([&] {
        State& s = sv(h);
        State& o = sv(outState);
        need(o.nq == s.nq && o.prec == s.prec && &o != &s, FSA_ERR_ARG, "output must be a distinct state of the same shape");
        need(numTerms >= 1 && coeffs && paulis, FSA_ERR_ARG, "null argument");
        check_paulis(paulis, numTerms * s.N);
        flush(s);
        o.pending.clear();
        hip_check(hipStreamSynchronize(o.stream), "hipStreamSynchronize");
        auto terms = pauli_terms(coeffs, numTerms, paulis, s.N, false);
        auto* dt = (const PauliTermDev*)s.ring.push(terms.data(), terms.size() * sizeof(PauliTermDev));
        if (s.prec == FSA_FP64) launch_apply_pauli_sum<double>(s.stream, (const cx<double>*)s.amps, (cx<double>*)o.amps, s.N, dt, numTerms);
        else launch_apply_pauli_sum<float>(s.stream, (const cx<float>*)s.amps, (cx<float>*)o.amps, s.N, dt, numTerms);
        hip_check(hipGetLastError(), "apply Pauli sum");
        hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    });
}

int fsa_statevector_innerProduct(fsa_state* ha, fsa_state* hb, double* out) {
    return guard([&] {
        State& a = sv(ha);
        State& b = sv(hb);
        need(a.nq == b.nq && a.prec == b.prec, FSA_ERR_ARG, "states differ in shape or precision");
        need(out != nullptr, FSA_ERR_ARG, "null output");
        flush(a);
        flush(b);
        hip_check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        double* ws = a.workspace(reduce_workspace_doubles(a.num()) + 8);
        if (a.prec == FSA_FP64) launch_inner<double>(a.stream, (const cx<double>*)a.amps, (const cx<double>*)b.amps, a.num(), ws + 8, ws);
        else launch_inner<float>(a.stream, (const cx<float>*)a.amps, (const cx<float>*)b.amps, a.num(), ws + 8, ws);
        hip_check(hipGetLastError(), "inner product");
        readback(a, ws, out, 2);
    });
}

}  // extern "C"
