// local_densitymatrix.hpp -- drop-in for the reference's src/local_densitymatrix.hpp:
// same names and signatures, executed by libfsa on the density matrix's
// device amplitudes (one rank: the local amplitudes are all of them).
#ifndef LOCAL_DENSITYMATRIX_HPP
#define LOCAL_DENSITYMATRIX_HPP

#include "local_statevector.hpp"

namespace fsa_dropin {
// the reference's depolarising coefficients are real in every call
inline double real_coeff(const Amp& c) {
    assert(c.imag() == 0);
    return c.real();
}
}  // namespace fsa_dropin

static void local_densitymatrix_oneQubitDephasing(DensityMatrix& rho, Nat qb, Real prob) {                  // :12-42
    fsa_dropin::one_rank(rho);
    fsa_check(fsa_densitymatrix_oneQubitDephasing(rho.handle, qb, prob), "oneQubitDephasing");
}

static void local_densitymatrix_twoQubitDephasing(DensityMatrix& rho, Nat qb1, Nat qb2, Real prob) {        // :45-60
    fsa_dropin::one_rank(rho);
    fsa_check(fsa_densitymatrix_twoQubitDephasing(rho.handle, qb1, qb2, prob), "twoQubitDephasing");
}

static void local_densitymatrix_oneQubitDepolarising(DensityMatrix& rho, Nat qb, Amp c1, Amp c2, Amp c3) {  // :63-81
    fsa_dropin::one_rank(rho);
    fsa_check(fsa_local_densitymatrix_oneQubitDepolarising(rho.handle, qb, fsa_dropin::real_coeff(c1),
                                                           fsa_dropin::real_coeff(c2), fsa_dropin::real_coeff(c3)),
              "oneQubitDepolarising");
}

static void local_densitymatrix_twoQubitDepolarising(DensityMatrix& rho, Nat q0, Nat q1, Nat q2, Nat q3, Amp c1,
                                                     Amp c2, Amp c3) {                                      // :83-108
    fsa_dropin::one_rank(rho);
    fsa_check(fsa_local_densitymatrix_twoQubitDepolarising(rho.handle, q0, q1, q2, q3, fsa_dropin::real_coeff(c1),
                                                           fsa_dropin::real_coeff(c2), fsa_dropin::real_coeff(c3)),
              "twoQubitDepolarising");
}

static void local_densitymatrix_damping(DensityMatrix& rho, Nat qb, Real prob) {                            // :111-131
    fsa_dropin::one_rank(rho);
    fsa_check(fsa_densitymatrix_damping(rho.handle, qb, prob), "damping");
}

// :134-164.  At one rank the reference pairs every traced row qubit t with
// its column qubit t + N (distributed_densitymatrix.hpp:365-369); that is the
// pairing the library traces.
static DensityMatrix local_densitymatrix_partialTrace(DensityMatrix& inRho, NatArray targs, NatArray pairTargs) {
    fsa_dropin::one_rank(inRho);
    assert(pairTargs.size() == targs.size());
    for (size_t i = 0; i < targs.size(); i++) assert(pairTargs[i] == targs[i] + inRho.numQubits);
    (void)pairTargs;
    fsa_state* out = nullptr;
    fsa_check(fsa_densitymatrix_partialTrace(inRho.handle, targs.data(), (unsigned)targs.size(), &out),
              "partialTrace");
    return DensityMatrix::adopt(out);
}

#endif  // LOCAL_DENSITYMATRIX_HPP
