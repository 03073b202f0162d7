/*
 * fsa.h -- C ABI of the MI355X full-state gate engine (libfsa.so).
 *
 * Every entry point is plain C: opaque handles, plain pointers and sizes,
 * int status codes (FSA_OK == 0).  On failure the message is available from
 * fsa_last_error() (per calling thread).  Amplitudes cross the boundary as
 * interleaved complex128 (re, im) doubles in the reference's layout:
 *   - qubit q is bit q of the amplitude index (reference src/types.hpp:26-37);
 *   - a density matrix of N qubits is its column-major 2N-qubit vector,
 *     amps[r + 2^N c] = rho[r][c] (reference tests/test_utilities.hpp:500);
 *   - a k-target gate matrix is 2^k x 2^k, row-major, matrix bit i acting on
 *     targets[i] in the caller's order (reference src/local_statevector.hpp:72-99);
 *   - Pauli codes I=0, X=1, Y=2, Z=3 (reference src/types.hpp:16-18).
 * Gate matrices are passed as interleaved complex128 row-major arrays
 * (2x2 -> 8 doubles).
 *
 * Execution: each state owns one HIP stream.  Operations validate their
 * arguments immediately and are queued on that stream (FSA_EXEC_EAGER: one
 * kernel per operation) or collected and applied at the next read/flush in
 * fused passes (FSA_EXEC_FUSED, the default): runs of operations whose
 * non-diagonal qubits fit one LDS tile are applied in a single HBM pass.
 * Reads (download, expectation values, partial trace) flush and synchronise.
 * A state is not thread-safe; distinct states may be used from distinct threads.
 */
#ifndef FSA_H
#define FSA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsa_state fsa_state;
typedef struct fsa_circuit fsa_circuit;

enum {
    FSA_OK = 0,
    FSA_ERR_ARG = 1,        /* null pointer, bad size, bad Pauli code, bad probability */
    FSA_ERR_QUBIT = 2,      /* qubit index out of range */
    FSA_ERR_DUPLICATE = 3,  /* repeated qubit among targets/controls */
    FSA_ERR_SIZE = 4,       /* too many qubits/targets for this build or device */
    FSA_ERR_OOM = 5,        /* device or pinned-host allocation failed */
    FSA_ERR_HIP = 6,        /* a HIP runtime call failed */
    FSA_ERR_KIND = 7,       /* statevector op on a density matrix or vice versa */
    FSA_ERR_INTERNAL = 8    /* an engine invariant failed (a bug in libfsa) */
};

enum { FSA_FP64 = 0, FSA_FP32 = 1 };
enum { FSA_EXEC_EAGER = 0, FSA_EXEC_FUSED = 1 };

const char* fsa_last_error(void);
int fsa_version(void);

/* ---------------------------------------------------------------------------
 * communication (reference src/communication.hpp:16-177).  On one GPU every
 * rank query answers rank 0 of 1 and the reductions are the identity.
 * ------------------------------------------------------------------------- */
int fsa_comm_init(void);                          /* comm_init        communication.hpp:16-21 */
int fsa_comm_end(void);                           /* comm_end         communication.hpp:24-27 */
unsigned fsa_comm_getRank(void);                  /* comm_getRank     communication.hpp:30-34 */
unsigned fsa_comm_getNumNodes(void);              /* comm_getNumNodes communication.hpp:37-41 */
int fsa_comm_synch(void);                         /* comm_synch       communication.hpp:44-47 (device-wide sync) */
int fsa_comm_reduceAmp(double* amp /* [2] */);    /* comm_reduceAmp   communication.hpp:172-177 (identity) */

/* ---------------------------------------------------------------------------
 * states (reference src/states.hpp:13-69).  numQubits is N for both kinds; a
 * density matrix holds 2^(2N) amplitudes.  New states are all-zero, like the
 * reference's zero-filled amps (states.hpp:46, 66).
 * ------------------------------------------------------------------------- */
int fsa_state_create(unsigned numQubits, int isDensityMatrix, int precision, fsa_state** out);
int fsa_state_destroy(fsa_state* s);
int fsa_state_info(const fsa_state* s, unsigned* numQubits, int* isDensityMatrix, int* precision,
                   uint64_t* numAmps);
int fsa_state_set_exec_mode(fsa_state* s, int mode);
int fsa_state_flush(fsa_state* s);          /* apply queued ops (asynchronous) */
int fsa_state_sync(fsa_state* s);           /* flush and wait for the state's stream */
int fsa_state_upload(fsa_state* s, const double* amps, uint64_t numAmps);
int fsa_state_download(fsa_state* s, double* amps, uint64_t numAmps);
int fsa_state_copy(fsa_state* dst, fsa_state* src);  /* same shape and precision */
int fsa_state_init_zero(fsa_state* s);      /* |0..0> (for a density matrix |0..0><0..0|) */
/* i.i.d. N(0,1) real and imaginary parts from a counter-based splitmix64 stream
 * + Box-Muller, normalised; the same law as the oracle's orc_init_gaussian */
int fsa_state_init_gaussian(fsa_state* s, uint64_t seed);
/* device address of the amplitudes (after a flush) and the state's stream,
 * for zero-copy interop (e.g. torch tensors) */
int fsa_state_device_ptr(fsa_state* s, void** ptr);
int fsa_state_stream(fsa_state* s, void** hipStream);
int fsa_state_norm2(fsa_state* s, double* out);   /* sum |amp|^2 in FP64 */
/* No reference counterpart: scale a statevector to unit norm, the norm summed
 * in FP64 on the device (an FP32 state drifts by about 1e-7 per gate).  Queued
 * on the state's stream after the pending operations; a zero state is left
 * as it is. */
int fsa_statevector_renormalise(fsa_state* psi);

/* ---------------------------------------------------------------------------
 * statevector operations (reference src/distributed_statevector.hpp)
 * ------------------------------------------------------------------------- */
int fsa_statevector_oneTargGate(fsa_state* psi, unsigned target, const double* gate);
    /* distributed_statevector_oneTargGate, distributed_statevector.hpp:18-40 */
int fsa_statevector_manyCtrlOneTargGate(fsa_state* psi, const unsigned* controls, unsigned numControls,
                                        unsigned target, const double* gate);
    /* distributed_statevector_manyCtrlOneTargGate, distributed_statevector.hpp:81-106 */
int fsa_statevector_swapGate(fsa_state* psi, unsigned qb1, unsigned qb2);
    /* distributed_statevector_swapGate, distributed_statevector.hpp:109-187 */
int fsa_statevector_manyTargGate(fsa_state* psi, const unsigned* targets, unsigned numTargets,
                                 const double* gate);
    /* distributed_statevector_manyTargGate, distributed_statevector.hpp:190-224 */
int fsa_statevector_pauliTensor(fsa_state* psi, const unsigned* targets, const unsigned* paulis,
                                unsigned numTargets);
    /* distributed_statevector_pauliTensor, distributed_statevector.hpp:279-284 */
int fsa_statevector_pauliGadget(fsa_state* psi, const unsigned* targets, const unsigned* paulis,
                                unsigned numTargets, double theta);
    /* distributed_statevector_pauliGadget, distributed_statevector.hpp:287-292: exp(+i theta P) */
int fsa_statevector_phaseGadget(fsa_state* psi, const unsigned* targets, unsigned numTargets,
                                double theta);
    /* distributed_statevector_phaseGadget, distributed_statevector.hpp:295-298 */

/* ---------------------------------------------------------------------------
 * density-matrix operations (reference src/distributed_densitymatrix.hpp)
 * ------------------------------------------------------------------------- */
int fsa_densitymatrix_manyTargGate(fsa_state* rho, const unsigned* targets, unsigned numTargets,
                                   const double* gate);                           /* :15-25 */
int fsa_densitymatrix_swapGate(fsa_state* rho, unsigned qb1, unsigned qb2);       /* :28-33 */
int fsa_densitymatrix_pauliTensor(fsa_state* rho, const unsigned* targets, const unsigned* paulis,
                                  unsigned numTargets);                           /* :36-50 */
int fsa_densitymatrix_pauliGadget(fsa_state* rho, const unsigned* targets, const unsigned* paulis,
                                  unsigned numTargets, double theta);             /* :53-64 */
int fsa_densitymatrix_phaseGadget(fsa_state* rho, const unsigned* targets, unsigned numTargets,
                                  double theta);                                  /* :67-76 */
int fsa_densitymatrix_krausMap(fsa_state* rho, const double* krausOps, unsigned numOps,
                               const unsigned* targets, unsigned numTargets);     /* :79-89 */
    /* krausOps: numOps consecutive 2^k x 2^k row-major complex matrices */
int fsa_densitymatrix_oneQubitDephasing(fsa_state* rho, unsigned qb, double prob);        /* :92-95 */
int fsa_densitymatrix_twoQubitDephasing(fsa_state* rho, unsigned qb1, unsigned qb2, double prob); /* :98-101 */
int fsa_densitymatrix_oneQubitDepolarising(fsa_state* rho, unsigned qb, double prob);     /* :104-143 */
int fsa_densitymatrix_twoQubitDepolarising(fsa_state* rho, unsigned qb1, unsigned qb2, double prob); /* :241-263 */
int fsa_densitymatrix_damping(fsa_state* rho, unsigned qb, double prob);                  /* :266-319 */
int fsa_densitymatrix_expecPauliString(fsa_state* rho, const double* coeffs, unsigned numTerms,
                                       const unsigned* allPaulis, double* out /* [2] */);
    /* :322-344; allPaulis is term-major, allPaulis[t*N + q] acts on qubit q */
int fsa_densitymatrix_partialTrace(fsa_state* rho, const unsigned* targets, unsigned numTargets,
                                   fsa_state** out);                              /* :347-406 */

/* ---------------------------------------------------------------------------
 * the reference's local subroutines whose arguments are not those of a public
 * operation (at one rank the local amplitudes are the whole state)
 * ------------------------------------------------------------------------- */
int fsa_local_statevector_pauliTensorOrGadget(fsa_state* psi, uint64_t maskXY, uint64_t maskYZ,
                                              const double* powI /* [2] */, const double* thisAmpFac /* [2] */,
                                              const double* otherAmpFac /* [2] */);
    /* local_statevector_pauliTensorOrGadget_subroutine, local_statevector.hpp:102-135:
     * psi <- thisAmpFac psi + otherAmpFac powI X^maskXY Z^maskYZ psi (Z applied first) */
int fsa_local_densitymatrix_oneQubitDepolarising(fsa_state* rho, unsigned qb, double c1, double c2, double c3);
    /* local_densitymatrix.hpp:63-81 with the caller's coefficients (real, as every caller passes) */
int fsa_local_densitymatrix_twoQubitDepolarising(fsa_state* rho, unsigned q0, unsigned q1, unsigned q2,
                                                 unsigned q3, double c1, double c2, double c3);
    /* local_densitymatrix.hpp:83-108 on any four distinct bits of the 2N-bit index: amplitudes
     * with bit q0 == bit q2 and bit q1 == bit q3 mix (c1 amp + c2 sum of the four), the others
     * scale by 1 + c3 */

/* No counterpart in the reference: fsa_densitymatrix_expecPauliString for
 * callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, same kernels, same result
 * bit for bit; its term list and partial sums live in a static 4 MiB device
 * array per GPU instead of memory the library allocates, and it calls no
 * hipMalloc/hipFree.  numTerms * (N + 6) must stay within about 0.5 M
 * (FSA_ERR_SIZE otherwise); the state must have no queued operations
 * (FSA_ERR_ARG: flush first, a flush may allocate).  Calls that overlap on one
 * GPU share that array and race. */
int fsa_densitymatrix_expecPauliString_noalloc(fsa_state* rho, const double* coeffs, unsigned numTerms,
                                               const unsigned* allPaulis, double* out /* [2] */);

/* The reference's local twoQubitDepolarising is not trace preserving (its
 * block-diagonal coefficient is 1-4p/5 where the textbook channel needs
 * 1-16p/15).  By default this library applies the textbook channel; passing 1
 * here reproduces the shipped arithmetic exactly (process-wide setting). */
int fsa_set_two_qubit_depolarising_as_shipped(int enable);

/* ---------------------------------------------------------------------------
 * extensions (no reference counterpart)
 * ------------------------------------------------------------------------- */
/* <psi| H |psi> for H = sum_t coeffs[t] P_t, paulis[t*n + q] on qubit q (statevector) */
int fsa_statevector_expecPauliSum(fsa_state* psi, const double* coeffs, unsigned numTerms,
                                  const unsigned* paulis, double* out /* [2] */);
/* out = H |psi> (out: a statevector of the same size and precision) */
int fsa_statevector_applyPauliSum(fsa_state* psi, const double* coeffs, unsigned numTerms,
                                  const unsigned* paulis, fsa_state* out);
/* <a|b> in FP64 */
int fsa_statevector_innerProduct(fsa_state* a, fsa_state* b, double* out /* [2] */);

/* Circuits: recorded operation lists applied to a state with a cached fused
 * schedule.  Operation arguments mean exactly what the statevector entry
 * points above mean.  Gadget angles are parameters that can be changed
 * between applications and differentiated by fsa_adjoint_gradient. */
int fsa_circuit_create(fsa_circuit** out);
int fsa_circuit_destroy(fsa_circuit* c);
int fsa_circuit_num_ops(const fsa_circuit* c, unsigned* out);
int fsa_circuit_num_params(const fsa_circuit* c, unsigned* out);
int fsa_circuit_oneTargGate(fsa_circuit* c, unsigned target, const double* gate);
int fsa_circuit_manyCtrlOneTargGate(fsa_circuit* c, const unsigned* controls, unsigned numControls,
                                    unsigned target, const double* gate);
int fsa_circuit_swapGate(fsa_circuit* c, unsigned qb1, unsigned qb2);
int fsa_circuit_manyTargGate(fsa_circuit* c, const unsigned* targets, unsigned numTargets,
                             const double* gate);
int fsa_circuit_pauliTensor(fsa_circuit* c, const unsigned* targets, const unsigned* paulis,
                            unsigned numTargets);
/* the gadget's angle becomes parameter number fsa_circuit_num_params() before the call */
int fsa_circuit_pauliGadget(fsa_circuit* c, const unsigned* targets, const unsigned* paulis,
                            unsigned numTargets, double theta);
int fsa_circuit_phaseGadget(fsa_circuit* c, const unsigned* targets, unsigned numTargets,
                            double theta);
int fsa_circuit_set_params(fsa_circuit* c, const double* thetas, unsigned numParams);
int fsa_circuit_apply(fsa_circuit* c, fsa_state* psi);       /* psi must be a statevector */
/* number of fused HBM passes the schedule for this state size uses */
int fsa_circuit_num_passes(fsa_circuit* c, unsigned numQubits, int precision, unsigned* out);

/* Adjoint-method gradient: psi0 is the input state (left unchanged), the
 * circuit is applied, and <H> and d<H>/d theta_p for every gadget parameter p
 * are returned (grads has fsa_circuit_num_params entries).  Reverse sweep by
 * inverse gates (no stored layers); inner products accumulate in FP64.  The
 * circuit keeps the sweep's two work states (twice psi0's size) for its next
 * call at the same size and precision; fsa_circuit_destroy frees them.  A
 * circuit must not be used from two threads at once. */
int fsa_adjoint_gradient(fsa_circuit* c, fsa_state* psi0, const double* coeffs, unsigned numTerms,
                         const unsigned* paulis, double* expec, double* grads);

/* Batches (no reference counterpart): many independent circuits, each on its
 * own statevector started from |0..0>, each with its own observable
 * H = sum_t coeffs[t] P_t (paulis[t*n + q] on qubit q).  fsa_batch_run
 * computes every <H> and every gadget gradient in ONE persistent kernel
 * launch (captured once as a HIP graph and replayed; new angles are one
 * copy).  Amplitudes are FP64 or FP32 (precision), all sums FP64.
 * fsa_batch_add copies the circuit's ops and current angles; the circuit may
 * be changed or destroyed afterwards.  Circuits may hold one-target,
 * controlled and dense (up to 5 targets) gates, swaps, Pauli tensors and
 * gadgets; every op must fit one tile (tile bits minus low bits qubits
 * outside the low qubits). */
typedef struct fsa_batch fsa_batch;
int fsa_batch_create(int precision, fsa_batch** out);
int fsa_batch_destroy(fsa_batch* b);
int fsa_batch_add(fsa_batch* b, const fsa_circuit* c, unsigned numQubits, const double* coeffs,
                  unsigned numTerms, const unsigned* paulis, unsigned* index);
int fsa_batch_info(const fsa_batch* b, unsigned* numCircuits, unsigned* numParams);
/* new angles for circuit `index` (num = that circuit's parameter count) */
int fsa_batch_set_params(fsa_batch* b, unsigned index, const double* thetas, unsigned num);
/* expecs[numCircuits]; grads[numParams], circuits in add order, each in its
 * own parameter order (either pointer may be NULL) */
int fsa_batch_run(fsa_batch* b, double* expecs, double* grads);
/* plan shape: workgroup size (256 or 512), tile bits (10..12 FP64, 10..13
 * FP32; states of at most that many qubits stay in LDS for the whole job) and
 * the number of low qubits always kept in a tile (HBM coalescing vs fusion) */
int fsa_batch_set_config(fsa_batch* b, int workgroupSize, int tileBits, int lowBits);
int fsa_batch_get_config(const fsa_batch* b, int* workgroupSize, int* tileBits, int* lowBits);
/* time every configuration on a spread sample of sampleCircuits circuits
 * (0 = all) and keep the fastest; bestMs = its time on the sample */
int fsa_batch_autotune(fsa_batch* b, unsigned sampleCircuits, double* bestMs);

/* Kernel timing: while enabled, every pass the library launches is bracketed
 * by HIP events on its stream; fsa_profile_read aggregates them per kernel
 * kind (count, total milliseconds).  Enabling clears previous records. */
typedef struct {
    char name[32];
    uint64_t count;
    double total_ms;
} fsa_profile_entry;
int fsa_profile_enable(int enable);
int fsa_profile_read(fsa_profile_entry* out, unsigned capacity, unsigned* numEntries);

#ifdef __cplusplus
}
#endif
#endif /* FSA_H */
