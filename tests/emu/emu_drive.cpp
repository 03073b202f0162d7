// Drives libfsa's real planner and kernel code on the CPU stand-in.
//
//   emu_drive circuit <file>              fused circuit (make_plan + execute: k_tile, streaming kernels);
//                                         EMU_EAGER=1 runs the ops through execute_now (as one list),
//                                         EMU_EAGER=2 one execute_now per op (each op alone: its
//                                         streaming kernel), EMU_FP32=1 on an FP32 state
//   emu_drive batch <file> TB BITS LOW P  batch (fsa_batch_* + k_batch), P = 0 FP64 / 1 FP32
//   emu_drive adjoint <file> P            fsa_adjoint_gradient from |0..0> (fused forward pass,
//                                         reverse-sweep kernels of kernels_reduce.hip)
//
// Input files are written by tests/test_emu_kernels.py.  Output: one line with
// the errors against the expected values in the file; exit status 0 unless
// something failed (AddressSanitizer reports end the process non-zero).
#include "hip/hip_runtime.h"
#include "../../distributed-full-state-algorithms_amd/csrc/engine.hpp"

#include <cmath>
#include <string>

using namespace fsa;

namespace {

struct Reader {
    FILE* f;
    template <class T>
    void get(T* p, size_t n) {
        if (n && fread(p, sizeof(T), n, f) != n) { std::fprintf(stderr, "short input\n"); std::exit(3); }
    }
    template <class T>
    T one() { T v; get(&v, 1); return v; }
};

// one op record: code, k, qubits[k], paulis[k] (gadgets, tensors), angle, matrix
int add_op(Reader& r, fsa_circuit* c, unsigned* numGadgets) {
    const uint32_t code = r.one<uint32_t>(), k = r.one<uint32_t>();
    std::vector<uint32_t> q(k), p;
    r.get(q.data(), k);
    if (code == 4 || code == 5) { p.resize(k); r.get(p.data(), k); }
    const double th = r.one<double>();
    std::vector<double> m;
    const unsigned d = code == 3 ? (1u << k) : 2u;
    if (code == 0 || code == 1 || code == 3) { m.resize(2 * d * d); r.get(m.data(), m.size()); }
    switch (code) {
    case 0: return fsa_circuit_oneTargGate(c, q[0], m.data());
    case 1: return fsa_circuit_manyCtrlOneTargGate(c, q.data(), k - 1, q[k - 1], m.data());
    case 2: return fsa_circuit_swapGate(c, q[0], q[1]);
    case 3: return fsa_circuit_manyTargGate(c, q.data(), k, m.data());
    case 4: return fsa_circuit_pauliTensor(c, q.data(), p.data(), k);
    case 5: ++*numGadgets; return fsa_circuit_pauliGadget(c, q.data(), p.data(), k, th);
    default: ++*numGadgets; return fsa_circuit_phaseGadget(c, q.data(), k, th);
    }
}

int run_circuit(const char* path) {
    Reader r{std::fopen(path, "rb")};
    if (!r.f) return 3;
    const unsigned n = r.one<uint32_t>(), L = r.one<uint32_t>();
    fsa_circuit* c;
    fsa_circuit_create(&c);
    unsigned g = 0;
    for (unsigned i = 0; i < L; i++)
        if (add_op(r, c, &g)) { std::printf("add failed: %s\n", fsa_last_error()); return 2; }
    const size_t N = 1ull << n;
    std::vector<cx<double>> v(N), want(N);
    r.get(v.data(), N);
    r.get(want.data(), N);
    std::fclose(r.f);
    const bool eager = std::getenv("EMU_EAGER") != nullptr, fp32 = std::getenv("EMU_FP32") != nullptr;
    State s;
    s.N = n;
    s.nq = n;
    s.prec = fp32 ? FSA_FP32 : FSA_FP64;
    s.mode = eager ? FSA_EXEC_EAGER : FSA_EXEC_FUSED;
    s.amps = std::malloc(N * (fp32 ? sizeof(cx<float>) : sizeof(cx<double>)));
    if (fp32)
        for (size_t i = 0; i < N; i++) static_cast<cx<float>*>(s.amps)[i] = mk<float>((float)v[i].x, (float)v[i].y);
    else
        std::memcpy(s.amps, v.data(), N * sizeof(cx<double>));
    Plan p = make_plan(c->c.ops, n, s.prec);
    size_t tiled = 0;
    for (const Stage& st : p.stages) tiled += st.tiled;
    profile_enable(true);
    if (eager && std::getenv("EMU_EAGER")[0] == '2') {
        for (const HOp& op : c->c.ops) execute_now(s, std::vector<HOp>{op});
    } else if (eager) {
        execute_now(s, c->c.ops);
    }
    if (!eager) execute(s, p);
    // launches per kernel family (the profiler's names), e.g. tile_pass_mfma=3
    for (const ProfEntry& e : profile_read()) std::printf("%s=%llu ", e.name.c_str(), (unsigned long long)e.count);
    profile_enable(false);
    double num = 0, den = 0;
    for (size_t i = 0; i < N; i++) {
        cx<double> a;
        if (fp32) {
            const cx<float> f = static_cast<cx<float>*>(s.amps)[i];
            a = mk<double>(f.x, f.y);
        } else {
            a = static_cast<cx<double>*>(s.amps)[i];
        }
        num += (a.x - want[i].x) * (a.x - want[i].x) + (a.y - want[i].y) * (a.y - want[i].y);
        den += want[i].x * want[i].x + want[i].y * want[i].y;
    }
    std::printf("passes=%zu tiled=%zu rel_err=%.3e\n", p.stages.size(), tiled, std::sqrt(num / den));
    std::free(s.amps);
    fsa_circuit_destroy(c);
    return 0;
}

int run_batch(const char* path, int tb, int bits, int low, int prec) {
    Reader r{std::fopen(path, "rb")};
    if (!r.f) return 3;
    const unsigned nj = r.one<uint32_t>();
    fsa_batch* b;
    fsa_batch_create(prec, &b);
    if (fsa_batch_set_config(b, tb, bits, low)) { std::printf("config: %s\n", fsa_last_error()); return 2; }
    std::vector<std::vector<double>> want;
    std::vector<fsa_circuit*> circuits;
    for (unsigned j = 0; j < nj; j++) {
        const unsigned n = r.one<uint32_t>(), L = r.one<uint32_t>(), T = r.one<uint32_t>();
        fsa_circuit* c;
        fsa_circuit_create(&c);
        unsigned g = 0;
        for (unsigned i = 0; i < L; i++)
            if (add_op(r, c, &g)) { std::printf("add failed: %s\n", fsa_last_error()); return 2; }
        std::vector<double> coeffs(T);
        std::vector<uint32_t> paulis((size_t)T * n);
        r.get(coeffs.data(), T);
        r.get(paulis.data(), paulis.size());
        std::vector<double> w(1 + g);
        r.get(w.data(), w.size());
        want.push_back(w);
        if (fsa_batch_add(b, c, n, coeffs.data(), T, paulis.data(), nullptr)) {
            std::printf("batch add: %s\n", fsa_last_error());
            return 2;
        }
        circuits.push_back(c);
    }
    std::fclose(r.f);
    unsigned nc = 0, npar = 0;
    fsa_batch_info(b, &nc, &npar);
    std::vector<double> e(nc), gr(npar);
    if (fsa_batch_run(b, e.data(), gr.data())) { std::printf("run: %s\n", fsa_last_error()); return 2; }
    double worstE = 0, num = 0, den = 0;
    size_t k = 0;
    for (unsigned j = 0; j < nc; j++) {
        worstE = std::max(worstE, std::fabs(e[j] - want[j][0]));
        for (size_t p = 1; p < want[j].size(); p++, k++) {
            num += (gr[k] - want[j][p]) * (gr[k] - want[j][p]);
            den += want[j][p] * want[j][p];
        }
    }
    std::printf("jobs=%u params=%u expec_abs_err=%.3e grad_rel_err=%.3e\n", nc, npar, worstE, std::sqrt(num / den));
    fsa_batch_destroy(b);
    for (fsa_circuit* c : circuits) fsa_circuit_destroy(c);
    return 0;
}

// one job in the batch record format: n, L, T, ops, coeffs, paulis, <H> and gradients
int run_adjoint(const char* path, int prec) {
    Reader r{std::fopen(path, "rb")};
    if (!r.f) return 3;
    const unsigned n = r.one<uint32_t>(), L = r.one<uint32_t>(), T = r.one<uint32_t>();
    fsa_circuit* c;
    fsa_circuit_create(&c);
    unsigned g = 0;
    for (unsigned i = 0; i < L; i++)
        if (add_op(r, c, &g)) { std::printf("add failed: %s\n", fsa_last_error()); return 2; }
    std::vector<double> coeffs(T), want(1 + g);
    std::vector<uint32_t> paulis((size_t)T * n);
    r.get(coeffs.data(), T);
    r.get(paulis.data(), paulis.size());
    r.get(want.data(), want.size());
    std::fclose(r.f);
    fsa_state st;
    State& s = st.s;
    s.N = n;
    s.nq = n;
    s.prec = prec;
    s.mode = FSA_EXEC_FUSED;
    hipMalloc(&s.amps, s.num() * s.amp_bytes());     // zero-filled
    if (prec == FSA_FP64) static_cast<cx<double>*>(s.amps)[0] = mk<double>(1, 0);
    else static_cast<cx<float>*>(s.amps)[0] = mk<float>(1, 0);
    std::vector<double> gr(g);
    double e = 0;
    profile_enable(true);
    if (fsa_adjoint_gradient(c, &st, coeffs.data(), T, paulis.data(), &e, gr.data())) {
        std::printf("adjoint: %s\n", fsa_last_error());
        return 2;
    }
    // launches per kernel family, e.g. adjoint_qrun=2
    for (const ProfEntry& p : profile_read()) std::printf("%s=%llu ", p.name.c_str(), (unsigned long long)p.count);
    profile_enable(false);
    double num = 0, den = 0;
    for (unsigned p = 0; p < g; p++) {
        num += (gr[p] - want[1 + p]) * (gr[p] - want[1 + p]);
        den += want[1 + p] * want[1 + p];
    }
    std::printf("params=%u expec_abs_err=%.3e grad_rel_err=%.3e\n", g, std::fabs(e - want[0]), std::sqrt(num / den));
    hipFree(s.amps);
    fsa_circuit_destroy(c);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 3 && std::string(argv[1]) == "circuit") return run_circuit(argv[2]);
    if (argc >= 7 && std::string(argv[1]) == "batch")
        return run_batch(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]));
    if (argc >= 4 && std::string(argv[1]) == "adjoint") return run_adjoint(argv[2], std::atoi(argv[3]));
    std::fprintf(stderr, "usage: emu_drive circuit FILE | batch FILE TB BITS LOW PREC | adjoint FILE PREC\n");
    return 1;
}
