// batch.cpp -- many independent circuits with adjoint gradients in one
// persistent kernel behind a captured HIP graph (no reference counterpart).
//
// Host side: every added circuit becomes one job with a forward plan, a
// reverse plan (inverse gates, gradient items in front of each gadget's
// inverse) and its observable's terms.  All plans of the batch live in one
// device blob uploaded once; angles live in a separate device array that a
// run refreshes with one copy, so the graph is captured once and replayed.
// The tile size, workgroup size and forced-low-qubit count of the plans are
// the batch's configuration, picked by fsa_batch_autotune.
#include "../../include/fsa.h"
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace fsa {

template <typename R>
int batch_max_bits();
template <typename R>
int batch_occupancy(int tb, int m);
template <typename R>
bool launch_batch(hipStream_t s, int tb, int m, unsigned grid, const BatchJob* jobs, unsigned numJobs,
                  const BatchStage* stages, const TileOp* items, const cx<double>* mats, const PauliTermDev* terms,
                  const double* params, cx<R>* arena, double* res, unsigned* counter);

constexpr unsigned MAX_STAGE_GRADS_HOST = 512;   // LDS gradient accumulators per pass (kernels_batch.hip)

struct BatchEntry {
    unsigned n = 0;
    std::vector<HOp> ops;
    unsigned numParams = 0;
    unsigned pbase = 0;                 // first parameter slot of this circuit
    std::vector<PauliTermDev> terms;
};

struct BatchConfig {
    int tb = 256;
    int m = 11;
    int low = 5;
};

struct Batch {
    int prec = FSA_FP64;
    std::vector<BatchEntry> entries;
    std::vector<double> thetas;         // all angles, entry order
    unsigned totalParams = 0;
    BatchConfig cfg;
    bool cfgSet = false;

    hipStream_t stream = nullptr;
    void* arena = nullptr;
    char* blob = nullptr;
    double* dparams = nullptr;
    double* dres = nullptr;
    unsigned* dcounter = nullptr;
    double* hres = nullptr;             // pinned readback
    double* hparams = nullptr;          // pinned upload
    size_t offJobs = 0, offStages = 0, offItems = 0, offMats = 0, offTerms = 0;
    size_t resDoubles = 0;
    unsigned grid = 0;
    bool built = false;
    bool paramsDirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~Batch() { release(); if (stream) (void)hipStreamDestroy(stream); }

    void release_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        release_graph();
        if (arena) (void)hipFree(arena);
        if (blob) (void)hipFree(blob);
        if (dparams) (void)hipFree(dparams);
        if (dres) (void)hipFree(dres);
        if (dcounter) (void)hipFree(dcounter);
        if (hres) (void)hipHostFree(hres);
        if (hparams) (void)hipHostFree(hparams);
        arena = nullptr; blob = nullptr; dparams = nullptr; dres = nullptr; dcounter = nullptr;
        hres = nullptr; hparams = nullptr;
        built = false;
    }
};

namespace {

void dev_alloc(void** p, size_t bytes, const char* what) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) {
        (void)hipGetLastError();
        fail(FSA_ERR_OOM, std::string("hipMalloc failed for ") + what);
    }
}

// GRAD marker in front of the inverse of a gadget: <lambda|P|phi> for its Pauli string
HOp grad_op(const HOp& g, unsigned slot) {
    HOp o;
    o.kind = OP_GRAD;
    if (g.kind == OP_PAULI) {
        o.mxy = g.mxy;
        o.myz = g.myz;
        o.pw = g.pw;
    } else {                          // phase gadget = exp(i theta Z..Z)
        o.mxy = 0;
        o.myz = g.pmask;
        o.pw = 1.0;
    }
    o.param = (int)slot;
    finish_op(o);
    return o;
}

template <typename R>
struct BlobBuilder {
    std::vector<BatchJob> jobs;
    std::vector<BatchStage> stages;
    std::vector<TileOp> items;
    std::vector<cx<double>> mats;      // FP64 in either precision: amplitudes combine in FP64
    std::vector<PauliTermDev> terms;

    // ops[] with per-op angle sign (+1 forward, -1 inverse; 0 for GRAD)
    void add_plan(const std::vector<HOp>& ops, const std::vector<int>& sign, unsigned n, SchedParams sp,
                  unsigned numJobs, bool rev, uint32_t& first, uint32_t& count) {
        std::vector<Stage> st = schedule(ops, n, sp);
        first = (uint32_t)stages.size();
        for (const Stage& s : st) {
            need(!s.tq.empty() && s.tq.size() == sp.m, FSA_ERR_SIZE,
                 "batched circuits need every op to fit one tile (at most tile bits - low bits high qubits)");
            int pos[64];
            for (int& p : pos) p = -1;
            BatchStage bs{};
            for (size_t i = 0; i < s.tq.size(); i++) {
                pos[s.tq[i]] = (int)i;
                bs.tileMask |= 1ULL << s.tq[i];
            }
            bs.m = (uint32_t)s.tq.size();
            bs.item0 = (uint32_t)items.size();
            uint32_t ngrad = 0;
            for (size_t ii = 0; ii < s.ops.size(); ii++) {
                const int oi = s.ops[ii];
                const HOp& op = ops[oi];
                TileOp t = translate_item<double>(op, pos, bs.tileMask, mats);
                if (op.kind == OP_GRAD && ngrad == MAX_STAGE_GRADS_HOST) {
                    // close the pass and continue the same ops in a new one on the same tiles
                    bs.nitems = (uint32_t)(items.size() - bs.item0);
                    stages.push_back(bs);
                    bs.item0 = (uint32_t)items.size();
                    ngrad = 0;
                }
                if (op.kind == OP_GRAD) {
                    t.p[0] = ngrad++;
                    t.mat_off = numJobs + (uint32_t)op.param;
                    t.param = NO_PARAM;
                    // the gadget's inverse right after it: one fused sweep (item field k = 1)
                    if (ii + 1 < s.ops.size()) {
                        const int ni = s.ops[ii + 1];
                        const HOp& nx = ops[ni];
                        if ((nx.kind == OP_PAULI || nx.kind == OP_PHASE) && nx.param == op.param && sign[ni] < 0)
                            t.k = 1;
                    }
                } else if (op.param >= 0 && (op.kind == OP_PAULI || op.kind == OP_PHASE)) {
                    t.param = (uint32_t)op.param;
                    t.f[4] = op.pw.real();
                    t.f[5] = op.pw.imag();
                    t.f[6] = (double)sign[oi];
                } else {
                    t.param = NO_PARAM;
                }
                items.push_back(t);
            }
            bs.nitems = (uint32_t)(items.size() - bs.item0);
            stages.push_back(bs);
        }
        count = (uint32_t)(stages.size() - first);
    }
};

template <typename R>
void build_impl(Batch& b) {
    const int M = std::min(b.cfg.m, batch_max_bits<R>());
    const unsigned numJobs = (unsigned)b.entries.size();
    BlobBuilder<R> bb;
    std::vector<double> cost(numJobs);
    uint64_t arenaElems = 0;
    for (unsigned j = 0; j < numJobs; j++) {
        const BatchEntry& e = b.entries[j];
        const bool resident = e.n <= (unsigned)M;
        SchedParams sp;
        sp.m = resident ? e.n : (unsigned)M;
        sp.low = resident ? e.n : std::min((unsigned)b.cfg.low, sp.m);
        // global parameter slots for this entry's gadgets
        std::vector<HOp> fwd = e.ops;
        for (HOp& op : fwd) if (op.param >= 0) op.param += (int)e.pbase;
        std::vector<int> fsign(fwd.size(), 1);
        std::vector<HOp> rev;
        std::vector<int> rsign;
        for (size_t k = fwd.size(); k-- > 0;) {
            if (fwd[k].param >= 0) {
                rev.push_back(grad_op(fwd[k], (unsigned)fwd[k].param));
                rsign.push_back(0);
            }
            rev.push_back(inverse(fwd[k]));
            rsign.push_back(-1);
        }
        BatchJob J{};
        J.n = e.n;
        J.resident = resident ? 1u : 0u;
        bb.add_plan(fwd, fsign, e.n, sp, numJobs, false, J.fwd0, J.nfwd);
        if (!resident && J.nfwd == 0) {
            // the first forward pass creates |0..0> in the arena, so there is always one
            BatchStage bs{};
            for (unsigned q = 0; q < sp.m; q++) bs.tileMask |= 1ULL << q;
            bs.m = sp.m;
            bs.item0 = (uint32_t)bb.items.size();
            J.fwd0 = (uint32_t)bb.stages.size();
            J.nfwd = 1;
            bb.stages.push_back(bs);
        }
        bb.add_plan(rev, rsign, e.n, sp, numJobs, true, J.rev0, J.nrev);
        J.term0 = (uint32_t)bb.terms.size();
        J.nterm = (uint32_t)e.terms.size();
        bb.terms.insert(bb.terms.end(), e.terms.begin(), e.terms.end());
        J.out = j;
        J.slot0 = numJobs + e.pbase;
        J.nslot = e.numParams;
        J.nrm = 2 * (numJobs + b.totalParams) + j;
        if (!resident) {
            J.phi = arenaElems;
            J.lam = arenaElems + (1ULL << e.n);
            arenaElems += 2ULL << e.n;
        }
        bb.jobs.push_back(J);
        // HBM passes (forward: phi; reverse: phi and lambda; observable: terms + 1)
        const double size = std::ldexp(1.0, (int)e.n);
        cost[j] = resident ? size * (double)(fwd.size() + 2 * rev.size()) / 64.0
                           : size * (2.0 * J.nfwd + 4.0 * J.nrev + J.nterm + 1.0);
    }
    // largest jobs first, so the last ones to start are short
    std::vector<unsigned> order(numJobs);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned c) { return cost[a] > cost[c]; });
    std::vector<BatchJob> sorted(numJobs);
    for (unsigned i = 0; i < numJobs; i++) sorted[i] = bb.jobs[order[i]];

    b.release();
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    b.offJobs = 0;
    b.offStages = align(sorted.size() * sizeof(BatchJob));
    b.offItems = align(b.offStages + bb.stages.size() * sizeof(BatchStage));
    b.offMats = align(b.offItems + bb.items.size() * sizeof(TileOp));
    b.offTerms = align(b.offMats + bb.mats.size() * sizeof(cx<double>));
    const size_t blobBytes = align(b.offTerms + bb.terms.size() * sizeof(PauliTermDev));
    std::vector<char> host(blobBytes, 0);
    std::memcpy(host.data() + b.offJobs, sorted.data(), sorted.size() * sizeof(BatchJob));
    std::memcpy(host.data() + b.offStages, bb.stages.data(), bb.stages.size() * sizeof(BatchStage));
    std::memcpy(host.data() + b.offItems, bb.items.data(), bb.items.size() * sizeof(TileOp));
    if (!bb.mats.empty()) std::memcpy(host.data() + b.offMats, bb.mats.data(), bb.mats.size() * sizeof(cx<double>));
    if (!bb.terms.empty())
        std::memcpy(host.data() + b.offTerms, bb.terms.data(), bb.terms.size() * sizeof(PauliTermDev));

    const size_t ampBytes = sizeof(cx<R>);
    dev_alloc(&b.arena, arenaElems * ampBytes, "the batch arena");
    dev_alloc((void**)&b.blob, blobBytes, "the batch plans");
    b.resDoubles = 2ULL * (numJobs + b.totalParams) + numJobs;   // <H>, gradients, <phi|phi>
    dev_alloc((void**)&b.dres, b.resDoubles * sizeof(double), "the batch results");
    dev_alloc((void**)&b.dparams, std::max<size_t>(1, b.totalParams) * sizeof(double), "the batch angles");
    dev_alloc((void**)&b.dcounter, 256, "the batch counter");
    hip_check(hipHostMalloc((void**)&b.hres, b.resDoubles * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&b.hparams, std::max<size_t>(1, b.totalParams) * sizeof(double),
                            hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipMemcpyAsync(b.blob, host.data(), blobBytes, hipMemcpyHostToDevice, b.stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");

    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const int occ = batch_occupancy<R>(b.cfg.tb, M);
    need(occ > 0, FSA_ERR_SIZE, "batch configuration does not fit the device");
    b.grid = (unsigned)std::min<uint64_t>((uint64_t)cus * occ, std::max(1u, numJobs));
    b.cfg.m = M;
    b.built = true;
    b.paramsDirty = true;
}

template <typename R>
void run_impl(Batch& b, double* expecs, double* grads) {
    if (!b.built) build_impl<R>(b);
    const unsigned numJobs = (unsigned)b.entries.size();
    if (b.paramsDirty && b.totalParams) {
        std::memcpy(b.hparams, b.thetas.data(), b.totalParams * sizeof(double));
        hip_check(hipMemcpyAsync(b.dparams, b.hparams, b.totalParams * sizeof(double), hipMemcpyHostToDevice,
                                 b.stream), "hipMemcpyAsync(angles)");
        b.paramsDirty = false;
    }
    if (!b.exec) {
        hip_check(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const bool ok = launch_batch<R>(b.stream, b.cfg.tb, b.cfg.m, b.grid,
                                        reinterpret_cast<const BatchJob*>(b.blob + b.offJobs), numJobs,
                                        reinterpret_cast<const BatchStage*>(b.blob + b.offStages),
                                        reinterpret_cast<const TileOp*>(b.blob + b.offItems),
                                        reinterpret_cast<const cx<double>*>(b.blob + b.offMats),
                                        reinterpret_cast<const PauliTermDev*>(b.blob + b.offTerms), b.dparams,
                                        static_cast<cx<R>*>(b.arena), b.dres, b.dcounter);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b.stream, &g);
        need(ok, FSA_ERR_SIZE, "no batch kernel for this configuration");
        hip_check(e, "hipStreamEndCapture");
        b.graph = g;
        hip_check(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    {
        ProfScope prof(b.prec == FSA_FP64 ? "batch_fp64" : "batch_fp32", b.stream);
        hip_check(hipGraphLaunch(b.exec, b.stream), "hipGraphLaunch");
    }
    hip_check(hipMemcpyAsync(b.hres, b.dres, b.resDoubles * sizeof(double), hipMemcpyDeviceToHost, b.stream),
              "hipMemcpyAsync(results)");
    hip_check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
    // FP32 renormalisation: the forward state of job j has <phi|phi> = 1 up to
    // rounding; <H> and every gradient term are quadratic in the state, so
    // dividing them by the FP64 norm <phi|phi> gives the values of the
    // normalised state
    const double* n2 = b.hres + 2 * (numJobs + b.totalParams);
    for (unsigned j = 0; j < numJobs; j++) {
        const BatchEntry& e = b.entries[j];
        const double sc = b.prec == FSA_FP32 ? 1.0 / n2[j] : 1.0;
        if (expecs) expecs[j] = b.hres[2 * j] * sc;
        if (grads)
            for (unsigned p = e.pbase; p < e.pbase + e.numParams; p++)
                grads[p] = -2.0 * b.hres[2 * (numJobs + p) + 1] * sc;
    }
}

void run(Batch& b, double* expecs, double* grads) {
    if (b.prec == FSA_FP64) run_impl<double>(b, expecs, grads);
    else run_impl<float>(b, expecs, grads);
}

}  // namespace
}  // namespace fsa

using namespace fsa;

struct fsa_batch {
    fsa::Batch b;
};

namespace {

Batch& bb(fsa_batch* h) {
    need(h != nullptr, FSA_ERR_ARG, "null batch");
    return h->b;
}

std::vector<PauliTermDev> sv_terms(unsigned n, const double* coeffs, unsigned T, const unsigned* paulis) {
    std::vector<PauliTermDev> terms(T);
    for (unsigned t = 0; t < T; t++) {
        uint64_t mxy = 0, myz = 0;
        cd ph = 1.0;
        for (unsigned q = 0; q < n; q++) {
            const unsigned code = paulis[(uint64_t)t * n + q];
            need(code <= 3, FSA_ERR_ARG, "Pauli codes are 0 (I), 1 (X), 2 (Y), 3 (Z)");
            if (code == 1 || code == 2) mxy |= 1ULL << q;
            if (code == 2 || code == 3) myz |= 1ULL << q;
            if (code == 2) ph *= cd(0, 1);
        }
        const cd f = coeffs[t] * ph;
        terms[t] = PauliTermDev{mxy, myz, f.real(), f.imag()};
    }
    return terms;
}

}  // namespace

extern "C" {

int fsa_batch_create(int precision, fsa_batch** out) {
    return guard([&] {
        need(out != nullptr, FSA_ERR_ARG, "null output handle");
        need(precision == FSA_FP64 || precision == FSA_FP32, FSA_ERR_ARG, "precision must be FSA_FP64 or FSA_FP32");
        auto* h = new fsa_batch();
        h->b.prec = precision;
        if (hipStreamCreateWithFlags(&h->b.stream, hipStreamNonBlocking) != hipSuccess) {
            delete h;
            fail(FSA_ERR_HIP, "hipStreamCreate failed");
        }
        h->b.cfg.m = precision == FSA_FP64 ? 11 : 12;
        *out = h;
    });
}

int fsa_batch_destroy(fsa_batch* h) {
    delete h;
    return FSA_OK;
}

int fsa_batch_add(fsa_batch* h, const fsa_circuit* circ, unsigned numQubits, const double* coeffs,
                  unsigned numTerms, const unsigned* paulis, unsigned* index) {
    return guard([&] {
        Batch& b = bb(h);
        need(circ != nullptr, FSA_ERR_ARG, "null circuit");
        need(numQubits >= 1 && numQubits <= 30, FSA_ERR_SIZE, "batched states hold 1 to 30 qubits");
        need(circ->c.maxQubit <= numQubits, FSA_ERR_QUBIT, "circuit uses a qubit beyond numQubits");
        need(numTerms == 0 || (coeffs && paulis), FSA_ERR_ARG, "null observable");
        for (const HOp& op : circ->c.ops)
            need(op.kind != OP_MATK || op.q.size() <= 5, FSA_ERR_SIZE, "batched dense gates have at most 5 targets");
        BatchEntry e;
        e.n = numQubits;
        e.ops = circ->c.ops;
        e.numParams = circ->c.numParams;
        e.pbase = b.totalParams;
        e.terms = sv_terms(numQubits, coeffs, numTerms, paulis);
        std::vector<double> th(e.numParams, 0.0);
        for (const HOp& op : e.ops)
            if (op.param >= 0) th[op.param] = op.theta;
        b.thetas.insert(b.thetas.end(), th.begin(), th.end());
        b.totalParams += e.numParams;
        if (index) *index = (unsigned)b.entries.size();
        b.entries.push_back(std::move(e));
        b.built = false;
    });
}

int fsa_batch_info(const fsa_batch* h, unsigned* numCircuits, unsigned* numParams) {
    return guard([&] {
        need(h != nullptr, FSA_ERR_ARG, "null batch");
        if (numCircuits) *numCircuits = (unsigned)h->b.entries.size();
        if (numParams) *numParams = h->b.totalParams;
    });
}

int fsa_batch_set_params(fsa_batch* h, unsigned index, const double* thetas, unsigned num) {
    return guard([&] {
        Batch& b = bb(h);
        need(index < b.entries.size(), FSA_ERR_ARG, "circuit index out of range");
        const BatchEntry& e = b.entries[index];
        need(num == e.numParams, FSA_ERR_ARG, "wrong number of angles for this circuit");
        need(num == 0 || thetas, FSA_ERR_ARG, "null angles");
        for (unsigned i = 0; i < num; i++) need(std::isfinite(thetas[i]), FSA_ERR_ARG, "angles must be finite");
        std::copy(thetas, thetas + num, b.thetas.begin() + e.pbase);
        b.paramsDirty = true;
    });
}

int fsa_batch_set_config(fsa_batch* h, int workgroupSize, int tileBits, int lowBits) {
    return guard([&] {
        Batch& b = bb(h);
        need(workgroupSize == 256 || workgroupSize == 512, FSA_ERR_ARG, "workgroup size must be 256 or 512");
        const int mmax = b.prec == FSA_FP64 ? batch_max_bits<double>() : batch_max_bits<float>();
        need(tileBits >= 10 && tileBits <= mmax, FSA_ERR_ARG, "tile bits out of range for this precision");
        need(lowBits >= 0 && lowBits <= tileBits - 3, FSA_ERR_ARG, "low bits must leave room for targets");
        b.cfg = BatchConfig{workgroupSize, tileBits, lowBits};
        b.cfgSet = true;
        b.built = false;
    });
}

int fsa_batch_get_config(const fsa_batch* h, int* workgroupSize, int* tileBits, int* lowBits) {
    return guard([&] {
        need(h != nullptr, FSA_ERR_ARG, "null batch");
        if (workgroupSize) *workgroupSize = h->b.cfg.tb;
        if (tileBits) *tileBits = h->b.cfg.m;
        if (lowBits) *lowBits = h->b.cfg.low;
    });
}

int fsa_batch_run(fsa_batch* h, double* expecs, double* grads) {
    return guard([&] {
        Batch& b = bb(h);
        need(!b.entries.empty(), FSA_ERR_ARG, "empty batch");
        run(b, expecs, grads);
    });
}

int fsa_batch_autotune(fsa_batch* h, unsigned sampleCircuits, double* bestMs) {
    return guard([&] {
        Batch& b = bb(h);
        need(!b.entries.empty(), FSA_ERR_ARG, "empty batch");
        // a sample of the batch, spread over it, timed for every candidate
        Batch s;
        s.prec = b.prec;
        hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
        const size_t N = b.entries.size();
        const size_t want = sampleCircuits == 0 ? N : std::min<size_t>(sampleCircuits, N);
        for (size_t i = 0; i < want; i++) {
            const BatchEntry& e = b.entries[(i * N) / want];
            BatchEntry c = e;
            c.pbase = s.totalParams;
            s.thetas.insert(s.thetas.end(), b.thetas.begin() + e.pbase, b.thetas.begin() + e.pbase + e.numParams);
            s.totalParams += e.numParams;
            s.entries.push_back(std::move(c));
        }
        const int mmax = b.prec == FSA_FP64 ? batch_max_bits<double>() : batch_max_bits<float>();
        std::vector<BatchConfig> cands;
        for (int m = 10; m <= mmax; m++)
            for (int tb : {256, 512}) cands.push_back(BatchConfig{tb, m, 5});
        auto time_cfg = [&](const BatchConfig& c) {
            s.cfg = c;
            s.built = false;
            run(s, nullptr, nullptr);           // build + capture + first replay
            const auto t0 = std::chrono::steady_clock::now();
            run(s, nullptr, nullptr);
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        };
        BatchConfig best = cands[0];
        double bestT = 1e300;
        for (const BatchConfig& c : cands) {
            const double t = time_cfg(c);
            if (t < bestT) { bestT = t; best = c; }
        }
        for (int low : {4, 6}) {                 // then the fusion depth at the best shape
            BatchConfig c = best;
            c.low = low;
            const double t = time_cfg(c);
            if (t < bestT) { bestT = t; best = c; }
        }
        b.cfg = best;
        b.cfgSet = true;
        b.built = false;
        if (bestMs) *bestMs = bestT;
    });
}

}  // extern "C"
