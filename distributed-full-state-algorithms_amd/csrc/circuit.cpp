// circuit.cpp -- recorded circuits (cached fused schedule) and the adjoint-method
// gradient (no reference counterpart).
//
// Adjoint method for U = U_K ... U_1 with gadgets U_k = exp(i theta_k P_k):
//   d<H>/d theta_k = 2 Re <psi| H U_K..U_{k+1} (i P_k) phi_k> = -2 Im <lambda_k| P_k |phi_k>,
// phi_k = U_k..U_1 psi0 and lambda_k = U_{k+1}^dag..U_K^dag H psi.  The reverse
// sweep starts from phi_K = psi and lambda_K = H psi and recomputes both by
// applying U_k^dag, so no intermediate state is stored.
#include "../../include/fsa.h"
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace fsa;

namespace fsa {

HOp inverse(const HOp& op) {
    HOp v = op;
    switch (op.kind) {
    case OP_MAT1:
    case OP_MATK: {
        uint64_t d = op.kind == OP_MAT1 ? 2 : (1ULL << op.q.size());
        for (uint64_t r = 0; r < d; r++)
            for (uint64_t k = 0; k < d; k++) v.mat[r * d + k] = std::conj(op.mat[k * d + r]);
        break;
    }
    case OP_PAULI:
        if (op.gadget) { v.theta = -op.theta; refresh_angle(v); }
        break;
    case OP_PHASE: v.theta = -op.theta; break;
    default: break;
    }
    return v;
}


}  // namespace fsa

namespace {


Circuit& cc(fsa_circuit* c) {
    need(c != nullptr, FSA_ERR_ARG, "null circuit");
    return c->c;
}

void check_list(const unsigned* q, unsigned num, const char* what) {
    need(num == 0 || q != nullptr, FSA_ERR_ARG, std::string("null ") + what);
    uint64_t seen = 0;
    for (unsigned i = 0; i < num; i++) {
        need(q[i] < 40, FSA_ERR_QUBIT, std::string(what) + " index out of range");
        need(!((seen >> q[i]) & 1), FSA_ERR_DUPLICATE, std::string("repeated qubit among ") + what);
        seen |= 1ULL << q[i];
    }
}

void add(Circuit& c, HOp op) {
    unsigned mx = 0;
    for (unsigned q = 0; q < 64; q++) if (((op.touch | op.ctrl) >> q) & 1) mx = q + 1;
    c.maxQubit = std::max(c.maxQubit, mx);
    c.ops.push_back(std::move(op));
    c.cachedPrec = -1;
}

void invalidate(Circuit& c) { c.cachedPrec = -1; }

std::vector<cd> mat_of(const double* g, uint64_t dim) {
    need(g != nullptr, FSA_ERR_ARG, "null gate matrix");
    std::vector<cd> m(dim * dim);
    for (uint64_t i = 0; i < dim * dim; i++) m[i] = cd(g[2 * i], g[2 * i + 1]);
    return m;
}

HOp pauli_op(const unsigned* t, const unsigned* p, unsigned num, bool gadget, double theta) {
    HOp o;
    o.kind = OP_PAULI;
    o.pw = 1.0;
    for (unsigned i = 0; i < num; i++) {
        need(p[i] <= 3, FSA_ERR_ARG, "Pauli codes are 0..3 (I, X, Y, Z)");
        if (p[i] == 2) o.pw *= cd(0, 1);
        if (p[i] == 1 || p[i] == 2) o.mxy |= 1ULL << t[i];
        if (p[i] == 2 || p[i] == 3) o.myz |= 1ULL << t[i];
    }
    o.gadget = gadget;
    o.theta = theta;
    refresh_angle(o);
    finish_op(o);
    return o;
}

const Plan& plan_for(Circuit& c, unsigned nq, int prec) {
    if (c.cachedNq != nq || c.cachedPrec != prec) {
        c.cached = make_plan(c.ops, nq, prec);
        c.cachedNq = nq;
        c.cachedPrec = prec;
    }
    return c.cached;
}


// a statevector work state on a stream it does not own
struct Scratch {
    State s;
    Scratch(unsigned n, int prec, hipStream_t stream) {
        s.N = n;
        s.nq = n;
        s.prec = prec;
        s.mode = FSA_EXEC_EAGER;
        s.stream = stream;
        s.ring.stream = stream;
        hip_check(hipMalloc(&s.amps, s.num() * s.amp_bytes()), "hipMalloc(adjoint scratch)");
    }
    ~Scratch() {
        (void)hipStreamSynchronize(s.stream);
        if (s.amps) (void)hipFree(s.amps);
        if (s.ws) (void)hipFree(s.ws);
    }
};

}  // namespace

namespace fsa {

// phi and lambda of the adjoint sweep on a stream of their own, with their
// parameter rings and workspace: allocated on the first call for a size and
// precision and reused by the next ones (a call allocates nothing then)
struct AdjointScratch {
    unsigned n;
    int prec;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;
    std::unique_ptr<Scratch> phi, lam;
    AdjointScratch(unsigned n_, int prec_) : n(n_), prec(prec_) {
        try {
            hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "hipEventCreate");
            phi = std::make_unique<Scratch>(n, prec, stream);
            lam = std::make_unique<Scratch>(n, prec, stream);
        } catch (...) {
            release();
            throw;
        }
    }
    ~AdjointScratch() { release(); }
    void release() {
        phi.reset();      // each synchronises the stream before freeing
        lam.reset();
        if (ready) (void)hipEventDestroy(ready);
        if (stream) (void)hipStreamDestroy(stream);
        ready = nullptr;
        stream = nullptr;
    }
};

}  // namespace fsa

namespace {

template <typename R>
void adjoint_impl(Circuit& c, State& psi0, const double* coeffs, unsigned T, const unsigned* paulis, double* expec,
                  double* grads) {
    const unsigned n = psi0.N;
    flush(psi0);
    if (!c.adj || c.adj->n != n || c.adj->prec != psi0.prec) {
        c.adj.reset();
        c.adj = std::make_shared<AdjointScratch>(n, psi0.prec);
    }
    AdjointScratch& A = *c.adj;
    hipStream_t st = A.stream;
    // psi0 is read below: after the work already queued on its own stream
    hip_check(hipEventRecord(A.ready, psi0.stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(st, A.ready, 0), "hipStreamWaitEvent");
    Scratch& phi = *A.phi;
    Scratch& lam = *A.lam;
    // results: slot 0 = <phi|lambda>, slot 1 + p = <lambda|P_p|phi>, then the
    // FP64 norms <psi0|psi0> and <phi|phi> (FP32 renormalisation), then the
    // gradient rows of the sweep
    const unsigned P = c.numParams;
    const size_t red = reduce_workspace_doubles(phi.s.num());
    const uint64_t stride = adjoint_row_stride(n);
    double* ws = phi.s.workspace(red + 2 * (P + 1) + 2 + stride * P + 8);
    double* res = ws + red;
    double* norms = res + 2 * (P + 1);
    double* rows = norms + 2;
    const bool renorm = psi0.prec == FSA_FP32;
    if (renorm) launch_norm2<R>(st, (const cx<R>*)psi0.amps, psi0.num(), ws, norms);
    hip_check(hipMemcpyAsync(phi.s.amps, psi0.amps, psi0.num() * psi0.amp_bytes(), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync");
    execute(phi.s, plan_for(c, n, psi0.prec));
    // a unitary circuit keeps the norm; FP32 rounding does not quite.  <H> and
    // every gradient term are quadratic in phi (lambda = H phi), so the
    // forward drift is undone by scaling them by <psi0|psi0> / <phi|phi> at
    // the end instead of rescaling the states
    if (renorm) launch_norm2<R>(st, (const cx<R>*)phi.s.amps, phi.s.num(), ws, norms + 1);

    std::vector<PauliTermDev> terms(T);
    for (unsigned t = 0; t < T; t++) {
        uint64_t mxy = 0, myz = 0;
        cd ph = 1.0;
        for (unsigned q = 0; q < n; q++) {
            unsigned code = paulis[(uint64_t)t * n + q];
            if (code == 1 || code == 2) mxy |= 1ULL << q;
            if (code == 2 || code == 3) myz |= 1ULL << q;
            if (code == 2) ph *= cd(0, 1);
        }
        cd f = coeffs[t] * ph;
        terms[t] = PauliTermDev{mxy, myz, f.real(), f.imag()};
    }
    auto* dt = (const PauliTermDev*)phi.s.ring.push(terms.data(), terms.size() * sizeof(PauliTermDev));
    auto* a_phi = (cx<R>*)phi.s.amps;
    auto* a_lam = (cx<R>*)lam.s.amps;
    launch_apply_pauli_sum<R>(st, a_phi, a_lam, n, dt, T);


    // Reverse sweep, op by op: for a gadget, <lambda|P_k|phi> and U_k^dag on phi
    // and on lambda in one pass over both states (k_adjoint_step); any other op
    // is undone on each state by its own kernel.
    // Runs of consecutive parametrised diagonal gadgets (phase gadgets, Z-only
    // Pauli gadgets) take one pass per ADJ_DIAG_MAX of them (k_adjoint_diag).
    // Every gadget's partial sums go to the row slot of its parameter; one
    // launch adds each slot's rows after the sweep (launch_grad_sums).
    std::vector<uint32_t> nbs(P, 0);
    launch_inner<R>(st, a_phi, a_lam, phi.s.num(), ws, res);
    auto diag_gadget = [](const HOp& o) {
        return o.param >= 0 && (o.kind == OP_PHASE || (o.kind == OP_PAULI && o.mxy == 0));
    };
    // Runs of consecutive parametrised one-qubit X/Y gadgets on distinct
    // qubits take one pass per ADJ_QRUN_MAX of them (k_adjoint_qrun).
    auto qubit_gadget = [](const HOp& o) {
        return o.param >= 0 && o.kind == OP_PAULI && __builtin_popcountll(o.mxy) == 1 && (o.myz & ~o.mxy) == 0;
    };
    // the diagonal run starting at op j (going backwards): returns its length
    auto build_diag = [&](size_t j0, AdjointDiagDev& d) {
        std::memset(&d, 0, sizeof(d));
        size_t j = j0 + 1;
        while (j-- > 0 && d.count < ADJ_DIAG_MAX && diag_gadget(c.ops[j])) {
            const HOp& g = c.ops[j];
            const HOp gi = inverse(g);
            const unsigned i = d.count++;
            cd e0, e1;
            if (g.kind == OP_PAULI) {              // th v + (+-fe) v, f = i^{#Y} (1: no Y)
                d.myz[i] = g.myz;
                d.fr[i] = g.pw.real(); d.fi[i] = g.pw.imag();
                e0 = gi.th + gi.fe;
                e1 = gi.th - gi.fe;
            } else {
                d.myz[i] = g.pmask;
                d.fr[i] = 1.0; d.fi[i] = 0.0;
                e0 = cd(std::cos(gi.theta), std::sin(gi.theta));
                e1 = cd(std::cos(gi.theta), -std::sin(gi.theta));
            }
            d.e0r[i] = e0.real(); d.e0i[i] = e0.imag();
            d.e1r[i] = e1.real(); d.e1i[i] = e1.imag();
            d.slot[i] = (uint32_t)g.param;
        }
        return d.count;
    };
    for (size_t k = c.ops.size(); k-- > 0;) {
        const HOp& op = c.ops[k];
        if (qubit_gadget(op) && k > 0 && qubit_gadget(c.ops[k - 1]) && c.ops[k - 1].mxy != op.mxy) {
            AdjointQubitRunDev d;
            std::memset(&d, 0, sizeof(d));
            uint64_t used = 0;
            size_t j = k + 1;
            while (j-- > 0 && d.count < ADJ_QRUN_MAX && qubit_gadget(c.ops[j]) && !(c.ops[j].mxy & used)) {
                const HOp& g = c.ops[j];
                const HOp gi = inverse(g);
                const unsigned i = d.count++;
                used |= g.mxy;
                d.q[i] = (uint32_t)__builtin_ctzll(g.mxy);
                d.y[i] = g.myz != 0;
                d.fr[i] = g.pw.real(); d.fi[i] = g.pw.imag();
                d.thr[i] = gi.th.real(); d.thi[i] = gi.th.imag();
                d.fer[i] = gi.fe.real(); d.fei[i] = gi.fe.imag();
                d.slot[i] = (uint32_t)g.param;
            }
            std::copy(d.q, d.q + d.count, d.sorted);
            std::sort(d.sorted, d.sorted + d.count);
            // a run on qubits below 6 takes the diagonal run after it (if any)
            // into the same pass
            AdjointDiagDev tail;
            unsigned nt = 0;
            if (used < 64 && n >= 6 && k >= d.count && diag_gadget(c.ops[k - d.count]))
                nt = build_diag(k - d.count, tail);
            ProfScope prof("adjoint_qrun", st);
            const unsigned nb = launch_adjoint_qrun<R>(st, a_phi, a_lam, n, d, rows, stride, nt ? &tail : nullptr);
            hip_check(hipGetLastError(), "adjoint qubit run launch");
            for (unsigned i = 0; i < d.count; i++) nbs[d.slot[i]] = nb;
            for (unsigned i = 0; i < nt; i++) nbs[tail.slot[i]] = nb;
            k -= d.count + nt - 1;  // the loop's k-- moves past the run and its tail
            continue;
        }
        if (diag_gadget(op) && k > 0 && diag_gadget(c.ops[k - 1])) {
            AdjointDiagDev d;
            build_diag(k, d);
            ProfScope prof("adjoint_diag", st);
            const unsigned nb = launch_adjoint_diag<R>(st, a_phi, a_lam, n, d, rows, stride);
            hip_check(hipGetLastError(), "adjoint diagonal run launch");
            for (unsigned i = 0; i < d.count; i++) nbs[d.slot[i]] = nb;
            k -= d.count - 1;       // the loop's k-- moves past the run
            continue;
        }
        const HOp inv = inverse(op);
        if (op.param >= 0) {
            AdjointStepDev d;
            std::memset(&d, 0, sizeof(d));
            if (op.kind == OP_PAULI) {
                d.mxy = op.mxy; d.myz = op.myz;
                d.fr = op.pw.real(); d.fi = op.pw.imag();
                d.thr = inv.th.real(); d.thi = inv.th.imag();
                d.fer = inv.fe.real(); d.fei = inv.fe.imag();
            } else {
                d.mxy = 0; d.myz = op.pmask;
                d.fr = 1.0; d.fi = 0.0;
                d.phase = 1;
                d.e0r = std::cos(inv.theta); d.e0i = std::sin(inv.theta);
                d.e1r = std::cos(inv.theta); d.e1i = -std::sin(inv.theta);
            }
            d.slot = (uint32_t)op.param;
            ProfScope prof("adjoint_step", st);
            nbs[op.param] = launch_adjoint_step<R>(st, a_phi, a_lam, n, d, rows, stride);
            hip_check(hipGetLastError(), "adjoint step launch");
            continue;
        }
        std::vector<HOp> invs{inv};
        execute_now(phi.s, invs);
        execute_now(lam.s, invs);
    }
    if (P) {
        auto* dnb = (const uint32_t*)phi.s.ring.push(nbs.data(), nbs.size() * sizeof(uint32_t));
        launch_grad_sums(st, rows, stride, dnb, P, res + 2);
        hip_check(hipGetLastError(), "gradient sums launch");
    }
    std::vector<double> host(2 * (P + 1) + 2);
    hip_check(hipMemcpyAsync(host.data(), res, host.size() * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    const double sc = renorm && host[2 * (P + 1) + 1] > 0 ? host[2 * (P + 1)] / host[2 * (P + 1) + 1] : 1.0;
    *expec = host[0] * sc;
    for (unsigned p = 0; p < P; p++) grads[p] = -2.0 * host[2 * (1 + p) + 1] * sc;
}

}  // namespace

extern "C" {

int fsa_circuit_create(fsa_circuit** out) {
    return guard([&] {
        need(out != nullptr, FSA_ERR_ARG, "null output handle");
        *out = new fsa_circuit();
    });
}

int fsa_circuit_destroy(fsa_circuit* c) {
    delete c;
    return FSA_OK;
}

int fsa_circuit_num_ops(const fsa_circuit* c, unsigned* out) {
    return guard([&] {
        need(c && out, FSA_ERR_ARG, "null argument");
        *out = (unsigned)c->c.ops.size();
    });
}

int fsa_circuit_num_params(const fsa_circuit* c, unsigned* out) {
    return guard([&] {
        need(c && out, FSA_ERR_ARG, "null argument");
        *out = c->c.numParams;
    });
}

int fsa_circuit_oneTargGate(fsa_circuit* h, unsigned target, const double* gate) {
    return guard([&] {
        Circuit& c = cc(h);
        check_list(&target, 1, "target");
        HOp o;
        o.kind = OP_MAT1;
        o.q = {target};
        o.mat = mat_of(gate, 2);
        finish_op(o);
        add(c, std::move(o));
    });
}

int fsa_circuit_manyCtrlOneTargGate(fsa_circuit* h, const unsigned* controls, unsigned numControls, unsigned target,
                                    const double* gate) {
    return guard([&] {
        Circuit& c = cc(h);
        need(numControls == 0 || controls, FSA_ERR_ARG, "null controls");
        std::vector<unsigned> all(controls, controls + numControls);
        all.push_back(target);
        check_list(all.data(), (unsigned)all.size(), "control/target");
        HOp o;
        o.kind = OP_MAT1;
        o.q = {target};
        for (unsigned i = 0; i < numControls; i++) o.ctrl |= 1ULL << controls[i];
        o.mat = mat_of(gate, 2);
        finish_op(o);
        add(c, std::move(o));
    });
}

int fsa_circuit_swapGate(fsa_circuit* h, unsigned qb1, unsigned qb2) {
    return guard([&] {
        Circuit& c = cc(h);
        need(qb1 < 40 && qb2 < 40, FSA_ERR_QUBIT, "qubit index out of range");
        if (qb1 == qb2) return;
        HOp o;
        o.kind = OP_SWAP;
        o.q = {std::min(qb1, qb2), std::max(qb1, qb2)};
        finish_op(o);
        add(c, std::move(o));
    });
}

int fsa_circuit_manyTargGate(fsa_circuit* h, const unsigned* targets, unsigned numTargets, const double* gate) {
    return guard([&] {
        Circuit& c = cc(h);
        need(numTargets >= 1, FSA_ERR_ARG, "manyTargGate needs at least one target");
        check_list(targets, numTargets, "target");
        need(numTargets <= 12, FSA_ERR_SIZE, "manyTargGate supports at most 12 targets");
        HOp o;
        o.kind = numTargets == 1 ? OP_MAT1 : OP_MATK;
        o.q.assign(targets, targets + numTargets);
        o.mat = mat_of(gate, 1ULL << numTargets);
        finish_op(o);
        add(c, std::move(o));
    });
}

int fsa_circuit_pauliTensor(fsa_circuit* h, const unsigned* targets, const unsigned* paulis, unsigned num) {
    return guard([&] {
        Circuit& c = cc(h);
        check_list(targets, num, "target");
        need(num == 0 || paulis, FSA_ERR_ARG, "null paulis");
        add(c, pauli_op(targets, paulis, num, false, 0.0));
    });
}

int fsa_circuit_pauliGadget(fsa_circuit* h, const unsigned* targets, const unsigned* paulis, unsigned num,
                            double theta) {
    return guard([&] {
        Circuit& c = cc(h);
        check_list(targets, num, "target");
        need(num == 0 || paulis, FSA_ERR_ARG, "null paulis");
        HOp o = pauli_op(targets, paulis, num, true, theta);
        o.param = (int)c.numParams++;
        add(c, std::move(o));
    });
}

int fsa_circuit_phaseGadget(fsa_circuit* h, const unsigned* targets, unsigned num, double theta) {
    return guard([&] {
        Circuit& c = cc(h);
        check_list(targets, num, "target");
        HOp o;
        o.kind = OP_PHASE;
        for (unsigned i = 0; i < num; i++) o.pmask |= 1ULL << targets[i];
        o.theta = theta;
        o.param = (int)c.numParams++;
        finish_op(o);
        add(c, std::move(o));
    });
}

int fsa_circuit_set_params(fsa_circuit* h, const double* thetas, unsigned numParams) {
    return guard([&] {
        Circuit& c = cc(h);
        need(numParams == c.numParams && (numParams == 0 || thetas), FSA_ERR_ARG, "parameter count mismatch");
        for (HOp& o : c.ops)
            if (o.param >= 0) {
                o.theta = thetas[o.param];
                refresh_angle(o);
            }
        for (HOp& o : c.cached.ops)
            if (o.param >= 0) {
                o.theta = thetas[o.param];
                refresh_angle(o);
            }
    });
}

int fsa_circuit_apply(fsa_circuit* h, fsa_state* psi) {
    return guard([&] {
        Circuit& c = cc(h);
        need(psi != nullptr, FSA_ERR_ARG, "null state");
        State& s = psi->s;
        need(!s.dm, FSA_ERR_KIND, "circuits act on statevectors");
        need(c.maxQubit <= s.N, FSA_ERR_QUBIT, "circuit uses a qubit the state does not have");
        flush(s);
        execute(s, plan_for(c, s.nq, s.prec));
    });
}

int fsa_circuit_num_passes(fsa_circuit* h, unsigned numQubits, int precision, unsigned* out) {
    return guard([&] {
        Circuit& c = cc(h);
        need(out != nullptr, FSA_ERR_ARG, "null output");
        need(precision == FSA_FP64 || precision == FSA_FP32, FSA_ERR_ARG, "bad precision");
        need(c.maxQubit <= numQubits, FSA_ERR_QUBIT, "circuit uses a qubit the state does not have");
        const auto& stages = plan_for(c, numQubits, precision).stages;
        unsigned n = 0;
        for (const Stage& st : stages) n += st.tiled ? 1 : (unsigned)st.ops.size();
        *out = n;
    });
}

int fsa_adjoint_gradient(fsa_circuit* h, fsa_state* psi0, const double* coeffs, unsigned numTerms,
                         const unsigned* paulis, double* expec, double* grads) {
    return guard([&] {
        Circuit& c = cc(h);
        need(psi0 != nullptr && expec != nullptr, FSA_ERR_ARG, "null argument");
        need(c.numParams == 0 || grads != nullptr, FSA_ERR_ARG, "null gradient output");
        need(numTerms >= 1 && coeffs && paulis, FSA_ERR_ARG, "the observable needs at least one term");
        State& s = psi0->s;
        need(!s.dm, FSA_ERR_KIND, "the adjoint gradient acts on statevectors");
        need(c.maxQubit <= s.N, FSA_ERR_QUBIT, "circuit uses a qubit the state does not have");
        for (uint64_t i = 0; i < (uint64_t)numTerms * s.N; i++)
            need(paulis[i] <= 3, FSA_ERR_ARG, "Pauli codes are 0..3 (I, X, Y, Z)");
        if (s.prec == FSA_FP64) adjoint_impl<double>(c, s, coeffs, numTerms, paulis, expec, grads);
        else adjoint_impl<float>(c, s, coeffs, numTerms, paulis, expec, grads);
    });
}

}  // extern "C"
