/*
 * ref_shim.cpp -- C entry points onto the UNMODIFIED reference headers
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Built by oracle/Makefile (target `ref`) with -I<reference>/src into
 * oracle/_ref/libfsa_ref.so.  It exists so the tests can run the reference's
 * own single-rank arithmetic on the same inputs as the oracle and pin the
 * oracle against it.  Only ops whose reference implementation is deterministic
 * and free of the setBit defect are meaningful comparisons; the others are
 * exported to measure that defect.  The reference is header-only and its state
 * classes call MPI, so MPI is initialised once (singleton, one rank).
 */
#include "types.hpp"
#include "states.hpp"
#include "distributed_statevector.hpp"
#include "distributed_densitymatrix.hpp"

#include <cstring>
#include <omp.h>

namespace {

void ensure_mpi() {
    static bool done = false;
    if (!done) { comm_init(); done = true; }
}

AmpMatrix to_matrix(const double* m, Index dim) {
    AmpMatrix out = getZeroMatrix(dim);
    for (Index r = 0; r < dim; r++)
        for (Index c = 0; c < dim; c++)
            out[r][c] = Amp(m[2 * (r * dim + c)], m[2 * (r * dim + c) + 1]);
    return out;
}

NatArray to_nats(const unsigned* v, unsigned num) { return NatArray(v, v + num); }

void load(StateVector& s, const double* amps) {
    std::memcpy((void*)s.amps.data(), amps, s.amps.size() * sizeof(Amp));
}
void store(StateVector& s, double* amps) {
    std::memcpy(amps, (const void*)s.amps.data(), s.amps.size() * sizeof(Amp));
}

}  // namespace

extern "C" {

int ref_set_threads(int n) { omp_set_num_threads(n); return omp_get_max_threads(); }

void ref_sv_oneTargGate(double* amps, unsigned n, unsigned t, const double* g) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_oneTargGate(s, t, to_matrix(g, 2)); store(s, amps);
}
void ref_sv_manyCtrlOneTargGate(double* amps, unsigned n, const unsigned* c, unsigned nc, unsigned t, const double* g) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_manyCtrlOneTargGate(s, to_nats(c, nc), t, to_matrix(g, 2)); store(s, amps);
}
void ref_sv_swapGate(double* amps, unsigned n, unsigned a, unsigned b) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_swapGate(s, a, b); store(s, amps);
}
void ref_sv_manyTargGate(double* amps, unsigned n, const unsigned* t, unsigned k, const double* g) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_manyTargGate(s, to_nats(t, k), to_matrix(g, 1ULL << k)); store(s, amps);
}
void ref_sv_pauliTensor(double* amps, unsigned n, const unsigned* t, const unsigned* p, unsigned k) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_pauliTensor(s, to_nats(t, k), to_nats(p, k)); store(s, amps);
}
void ref_sv_pauliGadget(double* amps, unsigned n, const unsigned* t, const unsigned* p, unsigned k, double th) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_pauliGadget(s, to_nats(t, k), to_nats(p, k), th); store(s, amps);
}
void ref_sv_phaseGadget(double* amps, unsigned n, const unsigned* t, unsigned k, double th) {
    ensure_mpi(); StateVector s(n); load(s, amps);
    distributed_statevector_phaseGadget(s, to_nats(t, k), th); store(s, amps);
}

void ref_dm_swapGate(double* amps, unsigned N, unsigned a, unsigned b) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_swapGate(r, a, b); store(r, amps);
}
void ref_dm_phaseGadget(double* amps, unsigned N, const unsigned* t, unsigned k, double th) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_phaseGadget(r, to_nats(t, k), th); store(r, amps);
}
void ref_dm_manyTargGate(double* amps, unsigned N, const unsigned* t, unsigned k, const double* g) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_manyTargGate(r, to_nats(t, k), to_matrix(g, 1ULL << k)); store(r, amps);
}
void ref_dm_pauliTensor(double* amps, unsigned N, const unsigned* t, const unsigned* p, unsigned k) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_pauliTensor(r, to_nats(t, k), to_nats(p, k)); store(r, amps);
}
void ref_dm_pauliGadget(double* amps, unsigned N, const unsigned* t, const unsigned* p, unsigned k, double th) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_pauliGadget(r, to_nats(t, k), to_nats(p, k), th); store(r, amps);
}
void ref_dm_oneQubitDephasing(double* amps, unsigned N, unsigned q, double p) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_oneQubitDephasing(r, q, p); store(r, amps);
}
void ref_dm_twoQubitDephasing(double* amps, unsigned N, unsigned a, unsigned b, double p) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_twoQubitDephasing(r, a, b, p); store(r, amps);
}
void ref_dm_oneQubitDepolarising(double* amps, unsigned N, unsigned q, double p) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_oneQubitDepolarising(r, q, p); store(r, amps);
}
void ref_dm_twoQubitDepolarising(double* amps, unsigned N, unsigned a, unsigned b, double p) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_twoQubitDepolarising(r, a, b, p); store(r, amps);
}
void ref_dm_damping(double* amps, unsigned N, unsigned q, double p) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    distributed_densitymatrix_damping(r, q, p); store(r, amps);
}
void ref_dm_expecPauliString(const double* amps, unsigned N, const double* coeffs, unsigned T,
                             const unsigned* paulis, double* out) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    Amp v = distributed_densitymatrix_expecPauliString(
        r, RealArray(coeffs, coeffs + T), to_nats(paulis, T * N));
    out[0] = v.real(); out[1] = v.imag();
}
void ref_dm_partialTrace(const double* amps, unsigned N, const unsigned* t, unsigned k, double* out) {
    ensure_mpi(); DensityMatrix r(N); load(r, amps);
    DensityMatrix o = distributed_densitymatrix_partialTrace(r, to_nats(t, k));
    store(o, out);
}

// ---- timing on a persistent state (CPU-baseline calibration) -------------------
// The entry points above copy the state in and out around each call; these
// keep one reference state alive and time only the reference call itself.
// op codes: 0 oneTargGate, 1 manyCtrlOneTargGate (2 controls), 2 swapGate,
// 3 phaseGadget; density matrices: 10 oneQubitDephasing, 11
// twoQubitDephasing, 12 oneQubitDepolarising, 13 damping, 14
// expecPauliString (one term per call).  q holds the qubits the op takes.
void* ref_time_state_new(unsigned n, int dm) {
    ensure_mpi();
    if (dm) return new DensityMatrix(n);
    return new StateVector(n);
}
void ref_time_state_free(void* s, int dm) {
    if (dm) delete static_cast<DensityMatrix*>(s);
    else delete static_cast<StateVector*>(s);
}
double ref_time_op(void* sp, int dm, int op, const unsigned* q, const double* g, double x, int reps) {
    StateVector& s = dm ? *static_cast<DensityMatrix*>(sp) : *static_cast<StateVector*>(sp);
    const AmpMatrix m = op <= 1 ? to_matrix(g, 2) : getZeroMatrix(2);
    const double t0 = omp_get_wtime();
    for (int r = 0; r < reps; r++) {
        switch (op) {
        case 0: distributed_statevector_oneTargGate(s, q[0], m); break;
        case 1: distributed_statevector_manyCtrlOneTargGate(s, NatArray{q[1], q[2]}, q[0], m); break;
        case 2: distributed_statevector_swapGate(s, q[0], q[1]); break;
        case 3: distributed_statevector_phaseGadget(s, NatArray(q, q + 4), x); break;
        case 10: distributed_densitymatrix_oneQubitDephasing(static_cast<DensityMatrix&>(s), q[0], x); break;
        case 11: distributed_densitymatrix_twoQubitDephasing(static_cast<DensityMatrix&>(s), q[0], q[1], x); break;
        case 12: distributed_densitymatrix_oneQubitDepolarising(static_cast<DensityMatrix&>(s), q[0], x); break;
        case 13: distributed_densitymatrix_damping(static_cast<DensityMatrix&>(s), q[0], x); break;
        case 14: {
            DensityMatrix& d = static_cast<DensityMatrix&>(s);
            NatArray codes(d.numQubits, 0);
            for (unsigned i = 0; i < 4 && i < d.numQubits; i++) codes[q[i]] = 1 + i % 3;
            (void)distributed_densitymatrix_expecPauliString(d, RealArray{x}, codes);
            break;
        }
        default: return -1.0;
        }
    }
    return (omp_get_wtime() - t0) / reps;
}

}  // extern "C"
