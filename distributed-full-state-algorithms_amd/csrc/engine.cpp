// engine.cpp -- fused-pass scheduler, parameter ring and op execution.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>

namespace fsa {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) fail(FSA_ERR_OOM, std::string(what) + ": out of device memory");
        fail(FSA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// ---- operation bookkeeping ---------------------------------------------------

void refresh_angle(HOp& op) {
    if (op.kind == OP_PAULI) {
        if (op.gadget) {
            op.th = cd(std::cos(op.theta), 0.0);
            op.fe = cd(0.0, std::sin(op.theta)) * op.pw;
        } else {
            op.th = 0.0;
            op.fe = cd(op.other_sign, 0.0) * op.pw;
        }
    }
}

void finish_op(HOp& op) {
    uint64_t m = 0;
    for (unsigned x : op.q) m |= 1ULL << x;
    switch (op.kind) {
    case OP_MAT1: op.nd = m; op.touch = m | op.ctrl; op.tileable = true; break;
    case OP_MATK: op.nd = m; op.touch = m; op.tileable = op.q.size() <= 5; break;
    case OP_SWAP: op.nd = m; op.touch = m; op.tileable = true; break;
    case OP_PAULI:
    case OP_GRAD: op.nd = op.mxy; op.touch = op.mxy | op.myz; op.tileable = true; break;
    case OP_PHASE: op.nd = 0; op.touch = op.pmask; op.tileable = true; break;
    case OP_DEPHASE1:
    case OP_DEPHASE2: op.nd = 0; op.touch = m; op.tileable = true; break;
    case OP_DEPOL1:
    case OP_DEPOL2:
    case OP_DAMP: op.nd = m; op.touch = m; op.tileable = true; break;
    }
}

// ---- scheduler ---------------------------------------------------------------

static unsigned env_u(const char* name, unsigned dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    return (unsigned)std::strtoul(v, nullptr, 10);
}

// Dense gates of at least this many targets (6-12) run as the MFMA GEMM of
// kernels_gemm.hip; FSA_GEMM_MINK=13 keeps them on the per-group kernel.  At 4
// and 5 targets the per-group kernel is faster (28 qubits, FP64: 1.70 vs 2.43
// ms at k = 4), so the GEMM starts at 6.
static unsigned gemm_min_k() { return std::max(6u, env_u("FSA_GEMM_MINK", 6)); }

// Tuning switches (diagnostics): FSA_MFMA=0 runs dense k=3..5 gates on the
// VALU form; FSA_SWEEPS=1 enables register sweeps for passes without dense
// gates.  Measured on a 30-qubit FP64 circuit of 1000 gates the defaults
// (MFMA on, sweeps off) are the fastest combination.
static bool env_flag(const char* name, bool dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    return v[0] != '0';
}

bool has_mfma_gates(const std::vector<HOp>& ops) {
    for (const HOp& op : ops)
        if (op.kind == OP_MATK && op.q.size() >= 3 && op.q.size() <= 5) return true;
    return false;
}

SchedParams sched_params(unsigned nq, int precision, bool mfmaGates) {
    // FP32 tiles hold 2^13 amplitudes (64 KiB, two workgroups per CU) where
    // dense gates of 3-5 targets need them: their MFMA form works on full
    // tiles.  Without such gates, 2^12 (32 KiB) lets four workgroups share a
    // CU, which hides more of each item's latency: the 24-qubit gadget
    // circuit's forward passes take 6.8 instead of 7.3 ms.
    const unsigned mmax = precision == FSA_FP64 ? (unsigned)tile_bits_max<double>()
                          : mfmaGates            ? (unsigned)tile_bits_max<float>()
                                                 : (unsigned)tile_bits_max<float>() - 1u;
    const unsigned mlim = precision == FSA_FP64 ? (unsigned)tile_bits_limit<double>() : (unsigned)tile_bits_limit<float>();
    unsigned m = std::min(env_u("FSA_TILE_BITS", mmax), mlim);
    // Qubits 0..low-1 are always tile qubits, so every HBM access of a pass is
    // a run of at least 2^low consecutive amplitudes: 128 bytes (one line) in
    // either precision.  Fewer forced qubits leave more tile slots for gates
    // (30-qubit FP64 circuit of 1000 gates: 173 passes at low 5, 138 at low 3,
    // 670.8 vs 626.5 gates/s); the scheduler still fills unused slots with the
    // next low qubits.
    const unsigned lowDefault = precision == FSA_FP64 ? 3u : 4u;
    unsigned low = std::min(env_u("FSA_TILE_LOW", lowDefault), m);
    if (nq <= m) { m = nq; low = nq; }
    return {m, low};
}

// one line per pass: "T|S <tile qubits> : kind[/k] ..." (diagnostics only)
static void dump_schedule(const char* path, const std::vector<HOp>& ops, const std::vector<Stage>& stages) {
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    for (const Stage& st : stages) {
        std::fprintf(f, "%s", st.tiled ? "T" : "S");
        for (unsigned q : st.tq) std::fprintf(f, " %u", q);
        std::fprintf(f, " :");
        for (int i : st.ops) {
            const HOp& o = ops[i];
            std::fprintf(f, " %u/%zu/%d", (unsigned)o.kind, o.q.size(), __builtin_popcountll(o.nd));
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

// Greedy pass builder.  An op joins the current pass when (a) its mixed qubits
// fit in the tile together with those already chosen and (b) it commutes with
// every op deferred so far in this scan (they either touch disjoint qubits or
// act diagonally on every shared one), so executing it early is exact.
// The tile of each pass grows from a seed: the low qubits alone, or the low
// qubits plus the mixed qubits of one of the first FSA_SCHED_SEEDS (16)
// fitting ops still waiting.  A seed scores the ops its scan takes plus the
// most that a scan of the ops it leaves takes from the best of their own
// FSA_SCHED_LOOK (16) seeds (0: this pass alone); scores count ops among the
// first FSA_SCHED_WINDOW (256) of a list, which bounds the planning time.  The
// first seed of the highest score wins.  On seeded 1000-gate circuits of 30
// qubits this gives 771 passes over six seeds of the generator against 823
// for the plain scan from the low qubits (128 against 138 for the default
// seed).
std::vector<Stage> schedule(const std::vector<HOp>& ops, unsigned nq, SchedParams sp) {
    std::vector<Stage> stages;
    const uint64_t lowMask = sp.low >= 64 ? ~0ULL : ((1ULL << sp.low) - 1ULL);
    const unsigned freeSlots = sp.m - sp.low;
    auto fits = [&](const HOp& op) {
        return op.tileable && (unsigned)__builtin_popcountll(op.nd & ~lowMask) <= freeSlots;
    };
    const unsigned seeds = env_u("FSA_SCHED_SEEDS", 16);
    const unsigned look = env_u("FSA_SCHED_LOOK", 16);
    const size_t window = std::max(1u, env_u("FSA_SCHED_WINDOW", 256));
    // one scan of the first `limit` ops of `list` from a seeded tile: the
    // tile it ends with, the ops it takes (in order) and the rest
    auto scan = [&](const std::vector<int>& list, size_t limit, uint64_t tile, std::vector<int>& tk,
                    std::vector<int>& df) {
        uint64_t blockedNd = 0, blockedAll = 0;
        tk.clear();
        df.clear();
        const size_t n = std::min(limit, list.size());
        for (size_t r = 0; r < n; r++) {
            const int i = list[r];
            const HOp& op = ops[i];
            bool take = false;
            if (fits(op) && !(op.nd & blockedAll) && !(op.touch & blockedNd)) {
                uint64_t nt = tile | op.nd;
                if ((unsigned)__builtin_popcountll(nt) <= sp.m) { tile = nt; take = true; }
            }
            if (take) {
                tk.push_back(i);
            } else {
                df.push_back(i);
                uint64_t nd = fits(op) ? op.nd : op.touch;
                blockedNd |= nd;
                blockedAll |= op.touch;
            }
        }
        return tile;
    };
    // the seeds of a list: the low qubits, then those plus the mixed qubits of
    // each of its first fitting ops that reach above them
    auto seedsOf = [&](const std::vector<int>& list, unsigned maxSeeds, std::vector<uint64_t>& out) {
        out.assign(1, lowMask);
        for (size_t r = 0; r < list.size() && out.size() <= maxSeeds; r++) {
            const HOp& op = ops[list[r]];
            if (fits(op) && (op.nd & ~lowMask) != 0) out.push_back(lowMask | op.nd);
        }
    };
    std::vector<int> remaining(ops.size());
    for (size_t i = 0; i < ops.size(); i++) remaining[i] = (int)i;
    std::vector<int> tk, df, tk2, df2;
    std::vector<uint64_t> s1, s2;
    while (!remaining.empty()) {
        if (!fits(ops[remaining[0]])) {
            Stage st;
            st.tiled = false;
            st.ops.push_back(remaining[0]);
            stages.push_back(st);
            remaining.erase(remaining.begin());
            continue;
        }
        uint64_t bestSeed = lowMask;
        size_t bestScore = 0;
        seedsOf(remaining, seeds, s1);
        if (s1.size() > 1) {
            for (size_t si = 0; si < s1.size(); si++) {
                scan(remaining, window, s1[si], tk, df);
                size_t score = tk.size(), best2 = 0;
                if (look && !df.empty()) {
                    seedsOf(df, look, s2);
                    for (uint64_t t2 : s2) {
                        scan(df, window, t2, tk2, df2);
                        best2 = std::max(best2, tk2.size());
                    }
                }
                score += best2;
                if (si == 0 || score > bestScore) {
                    bestScore = score;
                    bestSeed = s1[si];
                }
            }
        }
        std::vector<int> deferred;
        Stage st;
        uint64_t tile = scan(remaining, remaining.size(), bestSeed, st.ops, deferred);
        for (unsigned q = 0; q < nq && (unsigned)__builtin_popcountll(tile) < sp.m; q++) tile |= 1ULL << q;
        for (unsigned q = 0; q < nq; q++) if ((tile >> q) & 1) st.tq.push_back(q);
        st.tiled = st.ops.size() > 1;
        stages.push_back(st);
        remaining.swap(deferred);
    }
    if (const char* path = std::getenv("FSA_SCHEDULE_DUMP")) dump_schedule(path, ops, stages);
    return stages;
}

// ---- swap relabeling ------------------------------------------------------------

static uint64_t map_mask(uint64_t m, const unsigned* pos) {
    uint64_t r = 0;
    for (; m; m &= m - 1) r |= 1ULL << pos[__builtin_ctzll(m)];
    return r;
}

// Logical amplitude i lives at the physical index whose bit pos[q] is bit q
// of i.  A swap of logical qubits a, b only exchanges pos[a] and pos[b]; every
// other op is rewritten onto physical qubits.  At the end, physical swaps move
// each logical qubit home (at most nq - 1, chained through one position so
// they pack densely into tile passes).  Swaps move data exactly, so results
// are identical to executing the swaps in place.
std::vector<HOp> relabel_swaps(const std::vector<HOp>& ops, unsigned nq) {
    unsigned pos[64], who[64];
    for (unsigned q = 0; q < 64; q++) { pos[q] = q; who[q] = q; }
    std::vector<HOp> out;
    out.reserve(ops.size() + nq);
    for (const HOp& op : ops) {
        if (op.kind == OP_SWAP) {
            unsigned a = op.q[0], b = op.q[1];
            std::swap(pos[a], pos[b]);
            who[pos[a]] = a;
            who[pos[b]] = b;
            continue;
        }
        HOp o = op;
        for (unsigned& q : o.q) q = pos[q];
        o.ctrl = map_mask(o.ctrl, pos);
        o.mxy = map_mask(o.mxy, pos);
        o.myz = map_mask(o.myz, pos);
        o.pmask = map_mask(o.pmask, pos);
        finish_op(o);
        out.push_back(std::move(o));
    }
    for (unsigned p = 0; p < nq; p++) {
        while (who[p] != p) {
            unsigned l = who[p];           // logical l sits at physical p; its home is physical l
            unsigned lq = who[l];
            HOp sw;
            sw.kind = OP_SWAP;
            sw.q = {std::min(p, l), std::max(p, l)};
            finish_op(sw);
            out.push_back(std::move(sw));
            who[l] = l;
            pos[l] = l;
            who[p] = lq;
            pos[lq] = p;
        }
    }
    return out;
}

Plan make_plan(const std::vector<HOp>& ops, unsigned nq, int precision) {
    Plan p;
    p.ops = relabel_swaps(ops, nq);
    p.stages = schedule(p.ops, nq, sched_params(nq, precision, has_mfma_gates(p.ops)));
    return p;
}

// ---- parameter ring ----------------------------------------------------------

ParamRing::~ParamRing() {
    // the owner has synchronised the stream before destroying the ring
    for (auto& b : retired) {
        (void)hipHostFree(b.first);
        (void)hipFree(b.second);
    }
    if (host) (void)hipHostFree(host);
    if (dev) (void)hipFree(dev);
}

void ParamRing::release_retired() {
    for (auto& b : retired) {
        hip_check(hipHostFree(b.first), "hipHostFree");
        hip_check(hipFree(b.second), "hipFree");
    }
    retired.clear();
}

void* ParamRing::push(const void* data, size_t bytes) {
    size_t need = (bytes + 255) & ~(size_t)255;
    if (head + need > cap) {
        if (holds == 0 && need <= cap) {
            // every region handed out so far is dead once the stream is idle
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            release_retired();
            head = 0;
        } else {
            // regions handed out earlier may still be read by kernels not yet
            // launched: keep their buffer and continue in a larger one
            if (holds == 0) {
                if (stream) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
                release_retired();
                if (host) hip_check(hipHostFree(host), "hipHostFree");
                if (dev) hip_check(hipFree(dev), "hipFree");
            } else if (host) {
                retired.emplace_back(host, dev);
            }
            host = nullptr;
            dev = nullptr;
            size_t ncap = std::max<size_t>({need, 2 * cap, (size_t)1 << 20});
            hip_check(hipHostMalloc((void**)&host, ncap, hipHostMallocDefault), "hipHostMalloc(param ring)");
            hip_check(hipMalloc((void**)&dev, ncap), "hipMalloc(param ring)");
            cap = ncap;
            head = 0;
        }
    }
    std::memcpy(host + head, data, bytes);
    hip_check(hipMemcpyAsync(dev + head, host + head, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(params)");
    void* d = dev + head;
    head += need;
    return d;
}

double* State::workspace(size_t doubles) {
    if (doubles > wsDoubles) {
        if (ws) {
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            hip_check(hipFree(ws), "hipFree");
            ws = nullptr;
        }
        size_t n = std::max<size_t>(doubles, 8192);
        hip_check(hipMalloc((void**)&ws, n * sizeof(double)), "hipMalloc(workspace)");
        wsDoubles = n;
    }
    return ws;
}

// ---- kernel timing -------------------------------------------------------------

namespace {

struct ProfRecord {
    const char* name;
    hipEvent_t start, stop;
};
bool g_prof = false;
std::vector<ProfRecord> g_records;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t prof_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

const char* op_name(OpKind k) {
    switch (k) {
    case OP_MAT1: return "mat1";
    case OP_MATK: return "matk";
    case OP_SWAP: return "swap";
    case OP_PAULI: return "pauli";
    case OP_PHASE: return "phase";
    case OP_DEPHASE1: return "dephase1";
    case OP_DEPHASE2: return "dephase2";
    case OP_DEPOL1: return "depol1";
    case OP_DEPOL2: return "depol2";
    case OP_DAMP: return "damp";
    default: break;
    }
    return "op";
}

}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t s) : stream_(s) {
    if (!g_prof) return;
    name_ = name;
    start_ = prof_event();
    stop_ = prof_event();
    hip_check(hipEventRecord(start_, stream_), "hipEventRecord");
}

ProfScope::~ProfScope() {
    if (!name_) return;
    (void)hipEventRecord(stop_, stream_);
    g_records.push_back(ProfRecord{name_, start_, stop_});
}

void profile_enable(bool on) {
    for (auto& r : g_records) { g_event_pool.push_back(r.start); g_event_pool.push_back(r.stop); }
    g_records.clear();
    g_prof = on;
}

std::vector<ProfEntry> profile_read() {
    std::vector<ProfEntry> out;
    for (auto& r : g_records) {
        hip_check(hipEventSynchronize(r.stop), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, r.start, r.stop), "hipEventElapsedTime");
        auto it = std::find_if(out.begin(), out.end(), [&](const ProfEntry& e) { return e.name == r.name; });
        if (it == out.end()) { out.push_back(ProfEntry{r.name, 0, 0.0}); it = out.end() - 1; }
        it->count++;
        it->ms += ms;
    }
    return out;
}

// ---- execution -----------------------------------------------------------------

namespace {

template <typename R>
cx<R> to_cx(cd v) { return mk<R>((R)v.real(), (R)v.imag()); }

uint32_t chan_code(unsigned q, const int* pos) {
    return pos[q] >= 0 ? (0x100u | (uint32_t)pos[q]) : (uint32_t)q;
}

template <typename R>
TileOp translate(const HOp& op, const int* pos, uint64_t tileMask, std::vector<cx<R>>& mats, bool mfma) {
    TileOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = op.kind;
    t.param = op.param < 0 ? NO_PARAM : (uint32_t)op.param;
    auto local_mask = [&](uint64_t gm) {
        uint32_t lm = 0;
        for (uint64_t m = gm & tileMask; m; m &= m - 1) lm |= 1u << pos[__builtin_ctzll(m)];
        return lm;
    };
    switch (op.kind) {
    case OP_MAT1:
        t.p[0] = (uint32_t)pos[op.q[0]];
        t.lmask = local_mask(op.ctrl);
        t.gmask = op.ctrl & ~tileMask;
        for (int i = 0; i < 4; i++) { t.f[2 * i] = op.mat[i].real(); t.f[2 * i + 1] = op.mat[i].imag(); }
        break;
    case OP_MATK: {
        const unsigned k = (unsigned)op.q.size();
        t.k = k;
        for (size_t i = 0; i < op.q.size(); i++) t.p[i] = (uint32_t)pos[op.q[i]];
        t.mat_off = (uint32_t)mats.size();
        if (mfma && k >= 3) {
            // MFMA form: real 2D x 2D matrix [[Ur, -Ui], [Ui, Ur]], row-major, in
            // complex slots of the parameter buffer (two reals per slot)
            t.lmask2 = 1;
            t.lmask = MAT_UNPLACED;         // the pass gives it an LDS offset
            const uint64_t D = 1ULL << k, D2 = 2 * D;
            std::vector<R> real(D2 * D2);
            for (uint64_t r = 0; r < D; r++)
                for (uint64_t c = 0; c < D; c++) {
                    cd u = op.mat[r * D + c];
                    real[r * D2 + c] = (R)u.real();
                    real[r * D2 + D + c] = (R)-u.imag();
                    real[(D + r) * D2 + c] = (R)u.imag();
                    real[(D + r) * D2 + D + c] = (R)u.real();
                }
            for (uint64_t i = 0; i < real.size(); i += 2) mats.push_back(mk<R>(real[i], real[i + 1]));
        } else {
            for (const cd& v : op.mat) mats.push_back(to_cx<R>(v));
        }
        break;
    }
    case OP_SWAP:
        t.p[0] = (uint32_t)pos[op.q[0]];
        t.p[1] = (uint32_t)pos[op.q[1]];
        break;
    case OP_PAULI:
        t.lmask = local_mask(op.mxy);
        t.lmask2 = local_mask(op.myz);
        t.gmask = op.myz & ~tileMask;
        t.f[0] = op.th.real(); t.f[1] = op.th.imag();
        t.f[2] = op.fe.real(); t.f[3] = op.fe.imag();
        break;
    case OP_GRAD:                       // <lambda|P|phi> tile sum; the caller sets the result slot
        t.lmask = local_mask(op.mxy);
        t.lmask2 = local_mask(op.myz);
        t.gmask = op.myz & ~tileMask;
        t.f[0] = op.pw.real(); t.f[1] = op.pw.imag();
        break;
    case OP_PHASE:
        t.lmask = local_mask(op.pmask);
        t.gmask = op.pmask & ~tileMask;
        t.f[0] = std::cos(op.theta); t.f[1] = std::sin(op.theta);
        t.f[2] = std::cos(op.theta); t.f[3] = -std::sin(op.theta);
        break;
    case OP_DEPHASE1:
        t.p[0] = chan_code(op.q[0], pos);
        t.p[1] = chan_code(op.q[1], pos);
        t.f[0] = op.c[0];
        break;
    case OP_DEPHASE2:
        for (int i = 0; i < 4; i++) t.p[i] = chan_code(op.q[i], pos);
        t.f[0] = 1.0 + op.c[0];
        break;
    case OP_DEPOL1:
    case OP_DAMP:
        t.p[0] = (uint32_t)pos[op.q[0]];
        t.p[1] = (uint32_t)pos[op.q[1]];
        t.f[0] = op.c[0]; t.f[1] = op.c[1]; t.f[2] = op.c[2];
        break;
    case OP_DEPOL2:
        for (int i = 0; i < 4; i++) t.p[i] = (uint32_t)pos[op.q[i]];
        t.f[0] = op.c[0]; t.f[1] = op.c[1]; t.f[2] = 1.0 + op.c[2];
        break;
    }
    return t;
}

template <typename R>
void run_standalone(State& s, const HOp& op) {
    ProfScope prof(op_name(op.kind), s.stream);
    cx<R>* a = static_cast<cx<R>*>(s.amps);
    const unsigned nq = s.nq;
    const unsigned N = s.N;
    switch (op.kind) {
    case OP_MAT1: {
        Mat2<R> g;
        for (int i = 0; i < 4; i++) g.m[i] = to_cx<R>(op.mat[i]);
        launch_mat1<R>(s.stream, a, nq, op.q[0], op.ctrl, g);
        break;
    }
    case OP_MATK: {
        const unsigned k = (unsigned)op.q.size();
        const uint64_t D = 1ULL << k;
        if (k >= gemm_min_k() && gemm_fits(nq, k)) {
            const GemmGeom g = gemm_geom(nq, op.q.data(), k, sizeof(R));
            std::vector<GemmFrag<R>> frag(D * D);
            gemm_pack_matrix<R>(op.mat.data(), k, frag.data());
            auto* df = static_cast<const GemmFrag<R>*>(s.ring.push(frag.data(), frag.size() * sizeof(GemmFrag<R>)));
            const uint64_t se = gemm_scratch_elems(g, nq, sizeof(R));
            cx<R>* scratch = se ? reinterpret_cast<cx<R>*>(s.workspace((se * sizeof(cx<R>) + 7) / 8)) : nullptr;
            launch_matk_gemm<R>(s.stream, a, nq, g, df, scratch);
        } else if (k <= 5) {
            std::vector<cx<R>> m(D * D);
            for (uint64_t i = 0; i < D * D; i++) m[i] = to_cx<R>(op.mat[i]);
            auto* dm = static_cast<const cx<R>*>(s.ring.push(m.data(), m.size() * sizeof(cx<R>)));
            launch_matk<R>(s.stream, a, nq, op.q.data(), k, dm, nullptr);
        } else {
            // transposed matrix followed by the 2^k index offsets
            size_t matBytes = D * D * sizeof(cx<R>);
            std::vector<char> blob(matBytes + D * sizeof(uint64_t));
            cx<R>* mt = reinterpret_cast<cx<R>*>(blob.data());
            for (uint64_t r = 0; r < D; r++)
                for (uint64_t c = 0; c < D; c++) mt[c * D + r] = to_cx<R>(op.mat[r * D + c]);
            uint64_t* offs = reinterpret_cast<uint64_t*>(blob.data() + matBytes);
            for (uint64_t l = 0; l < D; l++) {
                uint64_t o = 0;
                for (unsigned i = 0; i < k; i++) if ((l >> i) & 1) o |= 1ULL << op.q[i];
                offs[l] = o;
            }
            auto* dmt = static_cast<const cx<R>*>(s.ring.push(blob.data(), blob.size()));
            launch_matk<R>(s.stream, a, nq, op.q.data(), k, nullptr, dmt);
        }
        break;
    }
    case OP_SWAP: launch_swap<R>(s.stream, a, nq, op.q[0], op.q[1]); break;
    case OP_PAULI: launch_pauli<R>(s.stream, a, nq, op.mxy, op.myz, to_cx<R>(op.th), to_cx<R>(op.fe)); break;
    case OP_PHASE:
        launch_phase<R>(s.stream, a, nq, op.pmask, mk<R>((R)std::cos(op.theta), (R)std::sin(op.theta)),
                        mk<R>((R)std::cos(op.theta), (R)-std::sin(op.theta)));
        break;
    case OP_DEPHASE1: launch_dephase1<R>(s.stream, a, N, op.q[0], (R)op.c[0]); break;
    case OP_DEPHASE2: launch_dephase2<R>(s.stream, a, N, op.q[0], op.q[2], (R)op.c[0]); break;
    case OP_DEPOL1: launch_depol1<R>(s.stream, a, N, op.q[0], (R)op.c[0], (R)op.c[1], (R)op.c[2]); break;
    case OP_DEPOL2:
        launch_depol2<R>(s.stream, a, nq, op.q.data(), (R)op.c[0], (R)op.c[1], (R)op.c[2]);
        break;
    case OP_DAMP: launch_damp<R>(s.stream, a, N, op.q[0], (R)op.c[0], (R)op.c[1], (R)op.c[2]); break;
    }
    hip_check(hipGetLastError(), "kernel launch");
}

// ---- sweeps: runs of ops applied to register groups inside a tile pass -----

bool sweepable(const HOp& op, unsigned W) {
    switch (op.kind) {
    case OP_MAT1: case OP_PHASE: case OP_SWAP: case OP_DEPHASE1: case OP_DEPHASE2: case OP_DEPOL1: case OP_DAMP:
        return true;
    case OP_PAULI: return (unsigned)__builtin_popcountll(op.mxy) <= W;
    // dense gates in registers keep inputs and outputs live: 2 targets for FP64
    // (W = 3), 3 for FP32 (W = 4); larger ones run as LDS items
    case OP_MATK: return op.q.size() >= 2 && op.q.size() <= W - 1;
    default: return false;
    }
}

uint32_t tile_local(uint64_t gm, const int* pos) {
    uint32_t lm = 0;
    for (; gm; gm &= gm - 1) {
        int p = pos[__builtin_ctzll(gm)];
        if (p >= 0) lm |= 1u << p;
    }
    return lm;
}

struct MaskSplit {
    uint32_t sweep = 0;   // sweep-local bits
    uint32_t tile = 0;    // tile-local bits outside the sweep
    uint64_t out = 0;     // global bits outside the tile
};

MaskSplit split_mask(uint64_t gm, const int* pos, const int* sl) {
    MaskSplit r;
    for (; gm; gm &= gm - 1) {
        unsigned q = (unsigned)__builtin_ctzll(gm);
        int lp = pos[q];
        if (lp < 0) r.out |= 1ULL << q;
        else if (sl[lp] >= 0) r.sweep |= 1u << sl[lp];
        else r.tile |= 1u << lp;
    }
    return r;
}

template <typename R>
TileOp translate_micro(const HOp& op, const int* pos, const int* sl, std::vector<cx<R>>& mats) {
    TileOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = op.kind;
    t.param = op.param < 0 ? NO_PARAM : (uint32_t)op.param;
    auto sw = [&](unsigned q) { return (uint32_t)sl[pos[q]]; };
    switch (op.kind) {
    case OP_MAT1: {
        t.p[0] = sw(op.q[0]);
        MaskSplit c = split_mask(op.ctrl, pos, sl);
        t.lmask = c.sweep; t.gq = c.tile; t.gmask = c.out;
        for (int i = 0; i < 4; i++) { t.f[2 * i] = op.mat[i].real(); t.f[2 * i + 1] = op.mat[i].imag(); }
        break;
    }
    case OP_PAULI: {
        t.lmask = split_mask(op.mxy, pos, sl).sweep;
        MaskSplit z = split_mask(op.myz, pos, sl);
        t.lmask2 = z.sweep; t.gq = z.tile; t.gmask = z.out;
        t.f[0] = op.th.real(); t.f[1] = op.th.imag();
        t.f[2] = op.fe.real(); t.f[3] = op.fe.imag();
        break;
    }
    case OP_PHASE: {
        MaskSplit z = split_mask(op.pmask, pos, sl);
        t.lmask = z.sweep; t.gq = z.tile; t.gmask = z.out;
        t.f[0] = std::cos(op.theta); t.f[1] = std::sin(op.theta);
        t.f[2] = std::cos(op.theta); t.f[3] = -std::sin(op.theta);
        break;
    }
    case OP_SWAP: {
        uint32_t a = sw(op.q[0]), b = sw(op.q[1]);
        t.p[0] = std::min(a, b); t.p[1] = std::max(a, b);
        break;
    }
    case OP_MATK: {
        // matrix bit i acts on sweep bit s_i; reorder so bit j acts on the j-th smallest
        const unsigned k = (unsigned)op.q.size();
        std::vector<unsigned> si(k), order(k);
        for (unsigned i = 0; i < k; i++) { si[i] = sw(op.q[i]); order[i] = i; }
        std::sort(order.begin(), order.end(), [&](unsigned x, unsigned y) { return si[x] < si[y]; });
        const uint64_t D = 1ULL << k;
        auto orig = [&](uint64_t rNew) {
            uint64_t r = 0;
            for (unsigned j = 0; j < k; j++) if ((rNew >> j) & 1) r |= 1ULL << order[j];
            return r;
        };
        t.k = k;
        for (unsigned j = 0; j < k; j++) t.p[j] = si[order[j]];
        t.mat_off = (uint32_t)mats.size();
        for (uint64_t r = 0; r < D; r++)
            for (uint64_t c = 0; c < D; c++) {
                cd v = op.mat[orig(r) * D + orig(c)];
                mats.push_back(mk<R>((R)v.real(), (R)v.imag()));
            }
        break;
    }
    case OP_DEPHASE1:
    case OP_DEPHASE2:
        for (size_t i = 0; i < op.q.size(); i++) {
            unsigned q = op.q[i];
            int lp = pos[q];
            t.p[i] = lp < 0 ? q : sl[lp] >= 0 ? (0x200u | (uint32_t)sl[lp]) : (0x100u | (uint32_t)lp);
        }
        t.f[0] = op.kind == OP_DEPHASE1 ? op.c[0] : 1.0 + op.c[0];
        break;
    case OP_DEPOL1:
    case OP_DAMP: {
        uint32_t a = sw(op.q[0]), b = sw(op.q[1]);
        t.p[0] = std::min(a, b); t.p[1] = std::max(a, b);
        t.f[0] = op.c[0]; t.f[1] = op.c[1]; t.f[2] = op.c[2];
        break;
    }
    default: break;
    }
    return t;
}

// A dephasing channel, or one or two one-mask diagonals (phase gadgets, Pauli
// gadgets and tensors of Z only; `second` may be null), as an OP_DIAG4 item:
// amplitude l of a tile with base index b is multiplied by f[(p1 << 1) | p2],
// p1 = parity of (l & lmask) ^ parity of (b & gmask), p2 likewise with lmask2
// and gq.  The factors are the ones the ops' own tile items multiply by (1
// where they leave an amplitude unchanged; two diagonals multiply), so both
// forms give the same results up to rounding.
TileOp translate_diag4(const HOp& op, const HOp* second, const int* pos, uint64_t tileMask) {
    TileOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = OP_DIAG4;
    t.param = NO_PARAM;
    auto split = [&](uint64_t gm, uint32_t& lm, uint64_t& out) {
        lm = 0;
        for (uint64_t m = gm & tileMask; m; m &= m - 1) lm |= 1u << pos[__builtin_ctzll(m)];
        out = gm & ~tileMask;
    };
    auto set = [&](int i, cd v) { t.f[2 * i] = v.real(); t.f[2 * i + 1] = v.imag(); };
    // a one-mask diagonal: its parity mask and the factors on even / odd parity
    auto one_mask = [](const HOp& d, uint64_t& mask, cd& even, cd& odd) {
        if (d.kind == OP_PHASE) {          // e^{+i theta} on even parity, e^{-i theta} on odd
            mask = d.pmask;
            even = cd(std::cos(d.theta), std::sin(d.theta));
            odd = cd(std::cos(d.theta), -std::sin(d.theta));
        } else {                           // Z-only Pauli item: th + fe on even parity, th - fe on odd
            mask = d.myz;
            even = d.th + d.fe;
            odd = d.th - d.fe;
        }
    };
    switch (op.kind) {
    case OP_PHASE:
    case OP_PAULI: {   // mask 1 from op, mask 2 from the second op (factors 1 without one)
        uint64_t m1 = 0, m2 = 0;
        cd e1, o1, e2(1, 0), o2(1, 0);
        one_mask(op, m1, e1, o1);
        if (second) one_mask(*second, m2, e2, o2);
        split(m1, t.lmask, t.gmask);
        split(m2, t.lmask2, t.gq);
        set(0, e1 * e2); set(1, e1 * o2); set(2, o1 * e2); set(3, o1 * o2);
        break;
    }
    case OP_DEPHASE1:   // scale where the bits of q and q + N differ
    case OP_DEPHASE2: { // ... where those of q1 / q1 + N or of q2 / q2 + N differ
        if (second) fail(FSA_ERR_INTERNAL, "a dephasing channel fills a pass-end item alone");
        split((1ULL << op.q[0]) | (1ULL << op.q[1]), t.lmask, t.gmask);
        if (op.kind == OP_DEPHASE2) split((1ULL << op.q[2]) | (1ULL << op.q[3]), t.lmask2, t.gq);
        const cd one(1, 0), fac(op.kind == OP_DEPHASE2 ? 1.0 + op.c[0] : op.c[0], 0);   // the tile items' factors
        set(0, one);
        set(1, op.kind == OP_DEPHASE2 ? fac : one);
        set(2, fac);
        set(3, fac);
        break;
    }
    default:
        fail(FSA_ERR_INTERNAL, "op has no diagonal pass-end form");
    }
    return t;
}

template <typename R>
void execute_impl(State& s, const std::vector<HOp>& ops, const std::vector<Stage>& stages) {
    // the pass items pushed below are read by tile passes launched after the
    // standalone ops in between, which push parameters of their own
    ParamRing::Hold hold(s.ring);
    constexpr unsigned W = (unsigned)sweep_bits<R>();
    constexpr unsigned FULL = (unsigned)tile_bits_max<R>();
    std::vector<TileOp> items, micro;
    std::vector<cx<R>> mats;
    std::vector<size_t> first(stages.size(), 0), count(stages.size(), 0);
    std::vector<uint32_t> flags(stages.size(), 0), matLds(stages.size(), 0);
    std::vector<TileGeom> geoms(stages.size());
    std::vector<RTilePass> rpass(stages.size());
    // FP64 passes over full tiles may run register-resident (k_rtile) when
    // every op has a register form.  Measured at 30 qubits, a remap costs about
    // 1 ms and an exchange between lanes about 0.5 ms, while a dense gate of 3-5
    // targets costs 1.2-1.5 ms less than in the LDS tile kernel; so by default
    // (FSA_RTILE unset) only passes with dense gates, no remap and at most two
    // lane exchanges run register-resident.  FSA_RTILE=1: every pass that can,
    // FSA_RTILE=0: none.
    const char* rtEnv = std::getenv("FSA_RTILE");
    const int rtMode = !(sizeof(R) == 8) ? 0 : (rtEnv && *rtEnv) ? (rtEnv[0] == '0' ? 0 : 2) : 1;
    const bool useRtile = rtMode != 0;
    for (size_t si = 0; si < stages.size(); si++) {
        const Stage& st = stages[si];
        if (!st.tiled) continue;
        int pos[64];
        for (int& p : pos) p = -1;
        TileGeom g;
        std::memset(&g, 0, sizeof(g));
        const unsigned m = (unsigned)st.tq.size();
        g.m = (uint8_t)m;
        for (size_t i = 0; i < st.tq.size(); i++) {
            pos[st.tq[i]] = (int)i;
            g.q[i] = (uint8_t)st.tq[i];
            g.tileMask |= 1ULL << st.tq[i];
        }
        geoms[si] = g;
        if constexpr (sizeof(R) == 8) {
            if (useRtile && m == (unsigned)tile_bits_max<R>()) {
                RTilePass rp;
                const size_t i0 = items.size(), m0 = mats.size();
                bool take = plan_rtile(ops, st, items, mats, rp) && rp.matBytes <= rtile_mat_lds_max();
                if (take && rtMode == 1) {
                    unsigned remaps = 0, lanes = 0, dense = 0;
                    for (size_t i = rp.first; i < rp.first + rp.count; i++) {
                        const TileOp& t = items[i];
                        if (t.kind == OP_REMAP) remaps++;
                        else if (t.kind == OP_MATK) dense++;
                        else if (t.kind == OP_MAT1 && t.p[0] < 9) lanes++;
                        else if (t.kind == OP_PAULI && (t.lmask & 63u)) lanes++;
                    }
                    take = remaps == 0 && dense > 0 && lanes <= 2;
                }
                if (take) {
                    rpass[si] = rp;
                    first[si] = rp.first;
                    count[si] = rp.count;
                    flags[si] = 32u;
                    continue;
                }
                items.resize(i0);
                mats.resize(m0);
            }
        }
        first[si] = items.size();
        // group consecutive sweepable ops whose mixed tile bits fit in W bits
        struct Group { bool sweep; uint32_t bits; std::vector<int> ops; };
        std::vector<Group> groups;
        const bool useSweeps = env_flag("FSA_SWEEPS", false);
        // passes with dense gates (k >= 3) or two-qubit depolarising run as
        // plain LDS items on the dense kernel variant; sweeps serve the others
        bool denseStage = false;
        for (int oi : st.ops) {
            const HOp& op = ops[oi];
            if ((op.kind == OP_MATK && !sweepable(op, W)) || op.kind == OP_DEPOL2) denseStage = true;
        }
        // the MFMA kernel variants carry no VALU dense gates: a pass uses one only
        // when every dense gate in it has the MFMA form (k = 3..5 on a full tile)
        const bool useMfma = env_flag("FSA_MFMA", true);
        bool stageMfma = useMfma && m >= FULL;        // full tiles of 2^12 (FP64) / 2^13 amplitudes or more
        bool anyMfma = false;
        size_t mfmaBytes = 0;
        for (int oi : st.ops) {
            const HOp& op = ops[oi];
            if (op.kind != OP_MATK) continue;
            if (op.q.size() >= 3 && op.q.size() <= 5) {
                anyMfma = true;
                mfmaBytes += (2 * sizeof(R)) << (2 * op.q.size());
            } else if (op.q.size() != 2) {
                stageMfma = false;   // two-target gates run on VALU in either variant
            }
        }
        stageMfma = stageMfma && anyMfma;
        // MFMA matrices sit in LDS (2 * 4^k reals each) when they fit in its area
        const bool matsInLds = stageMfma && mfmaBytes <= tile_mat_lds_bytes(sizeof(cx<R>) << m);
        // Diagonal ops that commute with every later op of the pass move to its
        // end (order kept among them); the kernel applies those in registers as
        // the tile leaves LDS, without an LDS round trip of their own.  An op
        // commutes with a later op when the later one mixes none of its qubits.
        std::vector<int> body, tail;
        // (not in FP32 MFMA passes: their kernels have no registers left for it)
        const bool drainDiag = env_flag("FSA_DRAIN_DIAG", true) && !(sizeof(R) == 4 && stageMfma);
        // one pass-end item holds a dephasing channel or up to two one-mask
        // diagonals (phase gadgets, Z-only Pauli gadgets and tensors)
        auto one_mask = [](const HOp& d) { return d.kind == OP_PHASE || (d.kind == OP_PAULI && d.mxy == 0); };
        for (size_t i = 0; i < st.ops.size(); i++) {
            const HOp& d = ops[st.ops[i]];
            bool drain = drainDiag &&
                         (tail.empty() ? one_mask(d) || d.kind == OP_DEPHASE1 || d.kind == OP_DEPHASE2
                                       : tail.size() < 2 * DRAIN_MAX && one_mask(d) && one_mask(ops[tail[0]]));
            for (size_t j = i + 1; drain && j < st.ops.size(); j++) drain = (ops[st.ops[j]].nd & d.touch) == 0;
            (drain ? tail : body).push_back(st.ops[i]);
        }
        for (int oi : body) {
            const HOp& op = ops[oi];
            if (useSweeps && !denseStage && m >= W && sweepable(op, W)) {
                uint32_t nd = tile_local(op.nd, pos);
                if (!groups.empty() && groups.back().sweep &&
                    (unsigned)__builtin_popcount(groups.back().bits | nd) <= W) {
                    groups.back().bits |= nd;
                    groups.back().ops.push_back(oi);
                } else {
                    groups.push_back(Group{true, nd, {oi}});
                }
            } else {
                groups.push_back(Group{false, 0, {oi}});
            }
        }
        for (Group& gr : groups) {
            if (!gr.sweep) {
                items.push_back(translate<R>(ops[gr.ops[0]], pos, g.tileMask, mats, stageMfma));
                continue;
            }
            // fill to W bits with the highest free tile bits (keeps the low, contiguous ones for lanes)
            for (int b = (int)m - 1; b >= 0 && (unsigned)__builtin_popcount(gr.bits) < W; b--) gr.bits |= 1u << b;
            int sl[32];
            for (int& x : sl) x = -1;
            TileOp it;
            std::memset(&it, 0, sizeof(it));
            it.kind = OP_SWEEP;
            unsigned j = 0;
            for (unsigned b = 0; b < m; b++)
                if ((gr.bits >> b) & 1) { it.p[j] = b; sl[b] = (int)j; j++; }
            it.k = (uint32_t)gr.ops.size();
            it.mat_off = (uint32_t)micro.size();
            it.param = NO_PARAM;
            for (int oi : gr.ops) micro.push_back(translate_micro<R>(ops[oi], pos, sl, mats));
            items.push_back(it);
        }
        if (!tail.empty())
            items.push_back(translate_diag4(ops[tail[0]], tail.size() > 1 ? &ops[tail[1]] : nullptr, pos, g.tileMask));
        flags[si] |= (tail.empty() ? 0u : 1u) << 8;
        count[si] = items.size() - first[si];
        // every MFMA matrix of the pass goes to LDS (the left half of the real
        // form, 2 * 4^k reals)
        if (matsInLds) {
            uint32_t used = 0;
            for (size_t i = first[si]; i < items.size(); i++) {
                TileOp& it = items[i];
                if (it.kind != OP_MATK || it.lmask2 != 1u) continue;
                const uint32_t need = (uint32_t)(2u * sizeof(R)) << (2 * it.k);
                if (used + need > tile_mat_lds_bytes(sizeof(cx<R>) << m))
                    fail(FSA_ERR_INTERNAL, "MFMA matrices of a tile pass exceed its LDS budget");
                it.lmask = used / (uint32_t)sizeof(R);
                used += need;
            }
            matLds[si] = used;
        } else if (stageMfma) {
            // VAR 3: each matrix is staged at offset 0 just before its item
            uint32_t most = 0;
            for (size_t i = first[si]; i < items.size(); i++) {
                TileOp& it = items[i];
                if (it.kind != OP_MATK || it.lmask2 != 1u) continue;
                it.lmask = 0;
                most = std::max(most, (uint32_t)(2u * sizeof(R)) << (2 * it.k));
            }
            if (most > tile_mat_lds_bytes(sizeof(cx<R>) << m))
                fail(FSA_ERR_INTERNAL, "an MFMA matrix exceeds the LDS area of a tile pass");
            matLds[si] = most;
        }
        // kernel variant: 1 for dense gates / two-qubit depolarising, 2 for MFMA
        // dense gates with matrices in LDS, 3 with matrices in HBM; bit 4 marks sweeps
        uint32_t var = 0;
        bool hasSweep = false;
        for (size_t i = first[si]; i < first[si] + count[si]; i++) {
            if (items[i].kind == OP_MATK) var = std::max(var, items[i].lmask2 == 1u ? 2u : 1u);
            if (items[i].kind == OP_DEPOL2) var = std::max(var, 1u);
            if (items[i].kind == OP_SWEEP) hasSweep = true;
        }
        if (var == 2u && !matsInLds) var = 3u;
        for (size_t i = first[si]; i < first[si] + count[si]; i++)
            if (items[i].kind == OP_MATK && (items[i].lmask2 == 1u) != (var >= 2u) && !(var >= 2u && items[i].k == 2))
                fail(FSA_ERR_INTERNAL, "tile pass mixes MFMA and VALU dense gates");
        flags[si] |= var << 2;
        if (hasSweep) flags[si] |= 16u;
    }
    const TileOp* ditems = nullptr;
    const TileOp* dmicro = nullptr;
    const cx<R>* dmats = nullptr;
    if (!items.empty()) {
        size_t itemBytes = items.size() * sizeof(TileOp);
        size_t microOff = (itemBytes + 255) & ~(size_t)255;
        size_t microBytes = micro.size() * sizeof(TileOp);
        size_t matOff = (microOff + microBytes + 255) & ~(size_t)255;
        std::vector<char> blob(matOff + mats.size() * sizeof(cx<R>));
        std::memcpy(blob.data(), items.data(), itemBytes);
        if (!micro.empty()) std::memcpy(blob.data() + microOff, micro.data(), microBytes);
        if (!mats.empty()) std::memcpy(blob.data() + matOff, mats.data(), mats.size() * sizeof(cx<R>));
        char* d = static_cast<char*>(s.ring.push(blob.data(), blob.size()));
        ditems = reinterpret_cast<const TileOp*>(d);
        dmicro = reinterpret_cast<const TileOp*>(d + microOff);
        dmats = reinterpret_cast<const cx<R>*>(d + matOff);
    }
    for (size_t si = 0; si < stages.size(); si++) {
        const Stage& st = stages[si];
        if (st.tiled && (flags[si] & 32u)) {
            if constexpr (sizeof(R) == 8) {
                bool dense = false;
                for (size_t i = first[si]; i < first[si] + count[si]; i++) dense = dense || items[i].kind == OP_MATK;
                ProfScope prof(dense ? "rtile_pass_mfma" : "rtile_pass_plain", s.stream);
                launch_rtile(s.stream, static_cast<cx<double>*>(s.amps), s.nq, geoms[si], ditems + first[si],
                             (unsigned)count[si], dmats, rpass[si].loadMap, rpass[si].storeMap, rpass[si].matBytes);
                hip_check(hipGetLastError(), "register tile kernel launch");
            }
        } else if (st.tiled) {
            // one profile name per kernel instantiation family, so the timings
            // line up with per-kernel profiler statistics
            const uint32_t var = (flags[si] >> 2) & 3u;
            ProfScope prof(var == 3 ? "tile_pass_mfma_staged" : var == 2 ? "tile_pass_mfma" : var == 1 ? "tile_pass_dense"
                           : (flags[si] & 16u) ? "tile_pass_sweep" : "tile_pass_plain", s.stream);
            launch_tile<R>(s.stream, static_cast<cx<R>*>(s.amps), s.nq, geoms[si], ditems + first[si],
                           (unsigned)count[si], dmicro, dmats, flags[si], matLds[si]);
            hip_check(hipGetLastError(), "tile kernel launch");
        } else {
            for (int oi : st.ops) run_standalone<R>(s, ops[oi]);
        }
    }
}

}  // namespace

template <typename R>
TileOp translate_item(const HOp& op, const int* pos, uint64_t tileMask, std::vector<cx<R>>& mats) {
    return translate<R>(op, pos, tileMask, mats, false);
}
template TileOp translate_item<double>(const HOp&, const int*, uint64_t, std::vector<cx<double>>&);
template TileOp translate_item<float>(const HOp&, const int*, uint64_t, std::vector<cx<float>>&);

void execute(State& s, const std::vector<HOp>& ops, const std::vector<Stage>& stages) {
    if (s.prec == FSA_FP64) execute_impl<double>(s, ops, stages);
    else execute_impl<float>(s, ops, stages);
}

void execute_now(State& s, const std::vector<HOp>& ops) {
    if (ops.size() == 1) {
        Stage st;
        st.tiled = false;
        st.ops.push_back(0);
        execute(s, ops, {st});
        return;
    }
    execute(s, ops, schedule(ops, s.nq, sched_params(s.nq, s.prec, has_mfma_gates(ops))));
}

void flush(State& s) {
    if (s.pending.empty()) return;
    std::vector<HOp> ops;
    ops.swap(s.pending);
    execute(s, make_plan(ops, s.nq, s.prec));
}

}  // namespace fsa
