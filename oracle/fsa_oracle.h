/*
 * fsa_oracle.h -- CPU oracle for the full-state gate path (TEST INFRASTRUCTURE ONLY).
 *
 * A clean-room, single-rank restatement in C99 + OpenMP of the reference's
 * local_statevector / local_densitymatrix loops and the single-rank branches of
 * its distributed_* dispatchers.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as the
 * timed CPU baseline.  The GPU library never calls it.
 *
 * Conventions (reference src/types.hpp:16-42, tests/test_utilities.hpp:500):
 *   - qubit q is bit q of the amplitude index; amplitudes are interleaved
 *     complex128 (double _Complex);
 *   - a k-target matrix is row-major, matrix row/col bit i <-> targets[i] in the
 *     caller's order;
 *   - a density matrix of N qubits is a 2N-qubit vector, column-major:
 *     amps[r + 2^N c] = rho[r][c]; row qubit q = bit q, column qubit q = bit q+N;
 *   - Pauli codes I=0, X=1, Y=2, Z=3.
 *
 * Oracle contract where the as-shipped reference is defective:
 *   D1  setBit (bit_maths.hpp:62-67) keeps only the addressed bit; the oracle
 *       uses the evident intent (number & ~(1<<b)) | (v<<b).
 *   D2  an all-Z Pauli string has numInnerIts = 0 (local_statevector.hpp:108) and
 *       leaves the state unchanged; the oracle applies exp(i theta Z..Z) (gadget)
 *       or the Z..Z sign pattern (tensor), as the reference's own tests intend.
 *   D3  local twoQubitDepolarising (local_densitymatrix.hpp:83-108) multiplies the
 *       trace by 1+4p/15; the oracle applies the textbook channel
 *       (1-p) rho + p/15 sum_{P != II} P rho P.  The as-shipped arithmetic is
 *       kept as orc_dm_twoQubitDepolarising_asShipped so the deviation can be
 *       measured.
 */
#ifndef FSA_ORACLE_H
#define FSA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* amplitudes are passed as interleaved (re, im) doubles */
typedef double orc_real;

/* ---- initial states ---- */
void orc_init_gaussian(orc_real* amps, unsigned numQubitsInVec, uint64_t seed);
double orc_norm2(const orc_real* amps, unsigned numQubitsInVec);

/* ---- statevector ops (reference src/distributed_statevector.hpp, local branch) ---- */
void orc_sv_oneTargGate(orc_real* amps, unsigned n, unsigned target, const orc_real* gate /*2x2*/);
void orc_sv_manyCtrlOneTargGate(orc_real* amps, unsigned n, const unsigned* ctrls, unsigned numCtrls,
                                unsigned target, const orc_real* gate);
void orc_sv_swapGate(orc_real* amps, unsigned n, unsigned qb1, unsigned qb2);
void orc_sv_manyTargGate(orc_real* amps, unsigned n, const unsigned* targets, unsigned numTargets,
                         const orc_real* gate /*2^k x 2^k*/);
void orc_sv_pauliTensor(orc_real* amps, unsigned n, const unsigned* targets, const unsigned* paulis,
                        unsigned numTargets);
void orc_sv_pauliGadget(orc_real* amps, unsigned n, const unsigned* targets, const unsigned* paulis,
                        unsigned numTargets, double theta);
void orc_sv_phaseGadget(orc_real* amps, unsigned n, const unsigned* targets, unsigned numTargets,
                        double theta);

/* ---- density-matrix ops (reference src/distributed_densitymatrix.hpp, 1-rank branch) ---- */
void orc_dm_manyTargGate(orc_real* amps, unsigned N, const unsigned* targets, unsigned numTargets,
                         const orc_real* gate);
void orc_dm_swapGate(orc_real* amps, unsigned N, unsigned qb1, unsigned qb2);
void orc_dm_pauliTensor(orc_real* amps, unsigned N, const unsigned* targets, const unsigned* paulis,
                        unsigned numTargets);
void orc_dm_pauliGadget(orc_real* amps, unsigned N, const unsigned* targets, const unsigned* paulis,
                        unsigned numTargets, double theta);
void orc_dm_phaseGadget(orc_real* amps, unsigned N, const unsigned* targets, unsigned numTargets,
                        double theta);
void orc_dm_krausMap(orc_real* amps, unsigned N, const orc_real* krausOps /*numOps x 2^k x 2^k*/,
                     unsigned numOps, const unsigned* targets, unsigned numTargets);
void orc_dm_oneQubitDephasing(orc_real* amps, unsigned N, unsigned qb, double prob);
void orc_dm_twoQubitDephasing(orc_real* amps, unsigned N, unsigned qb1, unsigned qb2, double prob);
void orc_dm_oneQubitDepolarising(orc_real* amps, unsigned N, unsigned qb, double prob);
void orc_dm_twoQubitDepolarising(orc_real* amps, unsigned N, unsigned qb1, unsigned qb2, double prob);
void orc_dm_twoQubitDepolarising_asShipped(orc_real* amps, unsigned N, unsigned qb1, unsigned qb2,
                                           double prob);
void orc_dm_damping(orc_real* amps, unsigned N, unsigned qb, double prob);
void orc_dm_expecPauliString(const orc_real* amps, unsigned N, const double* coeffs, unsigned numTerms,
                             const unsigned* allPaulis /*numTerms x N*/, orc_real* out /*2*/);
void orc_dm_partialTrace(const orc_real* inAmps, unsigned N, const unsigned* targets,
                         unsigned numTargets, orc_real* outAmps /*2^(2(N-k))*/);

/* ---- extensions with no reference counterpart (needed by the adjoint gradient) ---- */
/* out = sum_t coeffs[t] * P_t |in>, P_t from codes paulis[t*n + q] acting on qubit q */
void orc_sv_applyPauliSum(const orc_real* in, orc_real* out, unsigned n, const double* coeffs,
                          unsigned numTerms, const unsigned* paulis);
/* <a|b> */
void orc_sv_innerProduct(const orc_real* a, const orc_real* b, unsigned n, orc_real* out /*2*/);
/* <a| P |b> for a single Pauli string given as (targets, paulis) */
void orc_sv_pauliMatrixElement(const orc_real* a, const orc_real* b, unsigned n,
                               const unsigned* targets, const unsigned* paulis, unsigned numTargets,
                               orc_real* out /*2*/);

/* number of OpenMP threads the oracle will use */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
