// engine.hpp -- host side of libfsa: operation records, the fused-pass
// scheduler, device memory and the per-state parameter ring.
#pragma once

#include "../../include/fsa.h"
#include "launch.hpp"

#include <complex>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace fsa {

using cd = std::complex<double>;

// thrown inside the library, turned into a status code at the C boundary
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);

// One primitive operation on the 2^nq amplitude vector (a density-matrix API
// call becomes one or two of these on the 2N-qubit vector).
struct HOp {
    OpKind kind = OP_MAT1;
    std::vector<unsigned> q;        // MAT1: {target}; MATK: targets (caller order); SWAP: {a, b};
                                    // channels: {q, q+N} or {q1, q2, q1+N, q2+N}
    uint64_t ctrl = 0;              // MAT1: control mask
    std::vector<cd> mat;            // MAT1: 2x2, MATK: 2^k x 2^k (row-major)
    uint64_t mxy = 0, myz = 0;      // PAULI masks
    cd th = 0, fe = 0;              // PAULI: this-factor and other * i^{#Y}
    cd pw = 1;                      // PAULI: i^{#Y} (adjoint needs it separately)
    uint64_t pmask = 0;             // PHASE mask
    double theta = 0;               // PAULI gadget / PHASE angle
    bool gadget = false;            // PAULI: exp(i theta P) (true) or the tensor P (false)
    double other_sign = 1;          // PAULI tensor: -1 folds the odd-Y negation of the DM wrapper
    double c[3] = {0, 0, 0};        // channel constants
    int param = -1;                 // circuits: parameter index of a gadget angle

    uint64_t nd = 0;                // qubits the op mixes (must be tile qubits)
    uint64_t touch = 0;             // all qubits the op reads
    bool tileable = true;
};

// the inverse of a unitary op (conjugate transpose; gadget angles negated)
HOp inverse(const HOp& op);

// recompute PAULI factors / PHASE factors from theta (gadget parameters)
void refresh_angle(HOp& op);
void finish_op(HOp& op);            // fills nd / touch / tileable

struct Stage {
    bool tiled = true;
    std::vector<unsigned> tq;       // sorted tile qubits (tiled)
    std::vector<int> ops;           // op indices in execution order
};

struct SchedParams {
    unsigned m;      // tile bits
    unsigned low;    // low qubits always in the tile (coalescing)
};
// mfmaGates: the ops hold dense gates of 3-5 targets (FP32 then keeps the
// 2^13-amplitude tiles their MFMA form needs)
SchedParams sched_params(unsigned nq, int precision, bool mfmaGates = true);
bool has_mfma_gates(const std::vector<HOp>& ops);
std::vector<Stage> schedule(const std::vector<HOp>& ops, unsigned nq, SchedParams sp);

// A fused plan: swap gates become relabelings of the later ops' qubits, the
// accumulated qubit permutation is undone by swaps at the end, then passes.
struct Plan {
    std::vector<HOp> ops;
    std::vector<Stage> stages;
};
std::vector<HOp> relabel_swaps(const std::vector<HOp>& ops, unsigned nq);
Plan make_plan(const std::vector<HOp>& ops, unsigned nq, int precision);

// Pinned-host / device ring for per-launch parameters (matrices, tile ops).
// A region handed out by push stays valid until the next push that finds the
// ring full while no Hold is active: that push waits for the stream and starts
// over at offset 0.  While a Hold is active (a caller that pushes several
// regions before launching all their kernels), a full ring moves on to a new,
// larger buffer and keeps the old one alive until the next such restart.
struct ParamRing {
    hipStream_t stream = nullptr;
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0, head = 0;
    int holds = 0;
    std::vector<std::pair<char*, char*>> retired;   // (host, device) buffers still referenced
    ~ParamRing();
    void* push(const void* data, size_t bytes);   // returns device address, copy queued
    struct Hold {
        ParamRing& r;
        explicit Hold(ParamRing& ring) : r(ring) { r.holds++; }
        ~Hold() { r.holds--; }
        Hold(const Hold&) = delete;
        Hold& operator=(const Hold&) = delete;
    };

private:
    void release_retired();
};

struct State {
    unsigned N = 0;          // qubits of the user-facing object
    unsigned nq = 0;         // bits of the amplitude index (N or 2N)
    bool dm = false;
    int prec = FSA_FP64;
    int mode = FSA_EXEC_FUSED;
    void* amps = nullptr;
    hipStream_t stream = nullptr;
    ParamRing ring;
    double* ws = nullptr;    // reduction workspace (device)
    size_t wsDoubles = 0;
    double* hres = nullptr;  // pinned readback of reduction results
    std::vector<HOp> pending;

    uint64_t num() const { return 1ULL << nq; }
    size_t amp_bytes() const { return prec == FSA_FP64 ? 16 : 8; }
    double* workspace(size_t doubles);
};

struct AdjointScratch;   // circuit.cpp

struct Circuit {
    std::vector<HOp> ops;
    unsigned numParams = 0;
    unsigned maxQubit = 0;   // highest qubit index used + 1
    // cached plan per (nq, precision)
    unsigned cachedNq = 0;
    int cachedPrec = -1;
    Plan cached;
    // the adjoint gradient's work states (phi, lambda, their stream), kept for
    // the next call at the same size and precision; freed with the circuit
    std::shared_ptr<AdjointScratch> adj;
};

// run ops on the state's stream; eager: each op as its own pass(es)
void execute(State& s, const std::vector<HOp>& ops, const std::vector<Stage>& stages);
inline void execute(State& s, const Plan& p) { execute(s, p.ops, p.stages); }
void execute_now(State& s, const std::vector<HOp>& ops);
void flush(State& s);

// One stage op as a tile item (VALU form for dense gates); pos[q] = tile
// position of qubit q or -1; dense matrices are appended to mats.
template <typename R>
TileOp translate_item(const HOp& op, const int* pos, uint64_t tileMask, std::vector<cx<R>>& mats);

// A register-resident pass (k_rtile, FP64 tiles of 2^12): its items in slot
// space (with OP_REMAP items between layouts), the layouts of its load and
// store (4 bits per slot), and the LDS bytes of its MFMA matrices.
struct RTilePass {
    uint64_t loadMap = 0, storeMap = 0;
    uint32_t matBytes = 0;
    size_t first = 0, count = 0;
};
// appends the pass's items and matrices; false if an op of the stage has no
// register-resident form (the stage then runs on the LDS tile kernel)
bool plan_rtile(const std::vector<HOp>& ops, const Stage& st, std::vector<TileOp>& items,
                std::vector<cx<double>>& mats, RTilePass& out);

// brackets launches on `stream` with HIP events while profiling is enabled
class ProfScope {
public:
    ProfScope(const char* name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

private:
    const char* name_ = nullptr;
    hipEvent_t start_ = nullptr, stop_ = nullptr;
    hipStream_t stream_;
};

struct ProfEntry {
    std::string name;
    uint64_t count;
    double ms;
};
void profile_enable(bool on);
std::vector<ProfEntry> profile_read();

// per-thread message behind fsa_last_error
void set_last_error(const std::string& msg);

// run f, turning library errors into a status code
template <class F>
int guard(F&& f) {
    try {
        f();
        return FSA_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return FSA_ERR_OOM;
    } catch (...) {
        set_last_error("unexpected exception");
        return FSA_ERR_HIP;
    }
}

inline void need(bool ok, int code, const std::string& msg) {
    if (!ok) fail(code, msg);
}

}  // namespace fsa

struct fsa_state {
    fsa::State s;
};
struct fsa_circuit {
    fsa::Circuit c;
};
