// distributed_densitymatrix.hpp -- drop-in for the reference's
// src/distributed_densitymatrix.hpp: same names and signatures, bodies call
// the libfsa C ABI.  twoQubitDepolarising applies the textbook channel
// (fsa_set_two_qubit_depolarising_as_shipped(1) restores the shipped constants).
#ifndef DISTRIBUTED_DENSITYMATRIX_HPP
#define DISTRIBUTED_DENSITYMATRIX_HPP

#include "distributed_statevector.hpp"
#include "local_densitymatrix.hpp"

static void distributed_densitymatrix_manyTargGate(DensityMatrix& rho, NatArray targets, AmpMatrix gate) { // :15-25
    fsa_check(fsa_densitymatrix_manyTargGate(rho.handle, targets.data(), (unsigned)targets.size(),
                                             fsa_dropin::flat(gate).data()), "dm manyTargGate");
}

static void distributed_densitymatrix_swapGate(DensityMatrix& rho, Nat qb1, Nat qb2) {                       // :28-33
    fsa_check(fsa_densitymatrix_swapGate(rho.handle, qb1, qb2), "dm swapGate");
}

static void distributed_densitymatrix_pauliTensor(DensityMatrix& rho, NatArray targets, NatArray paulis) {  // :36-50
    fsa_check(fsa_densitymatrix_pauliTensor(rho.handle, targets.data(), paulis.data(), (unsigned)targets.size()),
              "dm pauliTensor");
}

static void distributed_densitymatrix_pauliGadget(DensityMatrix& rho, NatArray targets, NatArray paulis,
                                                  Real theta) {                                             // :53-64
    fsa_check(fsa_densitymatrix_pauliGadget(rho.handle, targets.data(), paulis.data(), (unsigned)targets.size(),
                                            theta), "dm pauliGadget");
}

static void distributed_densitymatrix_phaseGadget(DensityMatrix& rho, NatArray targets, Real theta) {       // :67-76
    fsa_check(fsa_densitymatrix_phaseGadget(rho.handle, targets.data(), (unsigned)targets.size(), theta),
              "dm phaseGadget");
}

static void distributed_densitymatrix_krausMap(DensityMatrix& rho, MatrixArray krausOps, NatArray targets) {  // :79-89
    std::vector<double> all;
    for (const AmpMatrix& K : krausOps) {
        std::vector<double> f = fsa_dropin::flat(K);
        all.insert(all.end(), f.begin(), f.end());
    }
    fsa_check(fsa_densitymatrix_krausMap(rho.handle, all.data(), (unsigned)krausOps.size(), targets.data(),
                                         (unsigned)targets.size()), "dm krausMap");
}

static void distributed_densitymatrix_oneQubitDephasing(DensityMatrix& rho, Nat qb, Real prob) {             // :92-95
    fsa_check(fsa_densitymatrix_oneQubitDephasing(rho.handle, qb, prob), "oneQubitDephasing");
}

static void distributed_densitymatrix_twoQubitDephasing(DensityMatrix& rho, Nat qb1, Nat qb2, Real prob) {  // :98-101
    fsa_check(fsa_densitymatrix_twoQubitDephasing(rho.handle, qb1, qb2, prob), "twoQubitDephasing");
}

static void distributed_densitymatrix_oneQubitDepolarising(DensityMatrix& rho, Nat qb, Real prob) {         // :104-143
    fsa_check(fsa_densitymatrix_oneQubitDepolarising(rho.handle, qb, prob), "oneQubitDepolarising");
}

static void distributed_densitymatrix_twoQubitDepolarising(DensityMatrix& rho, Nat qb1, Nat qb2, Real prob) { // :241-263
    fsa_check(fsa_densitymatrix_twoQubitDepolarising(rho.handle, qb1, qb2, prob), "twoQubitDepolarising");
}

static void distributed_densitymatrix_damping(DensityMatrix& rho, Nat qb, Real prob) {                      // :266-319
    fsa_check(fsa_densitymatrix_damping(rho.handle, qb, prob), "damping");
}

static Amp distributed_densitymatrix_expecPauliString(DensityMatrix& rho, RealArray coeffs,
                                                      NatArray allPaulis) {                                 // :322-344
    double out[2];
    fsa_check(fsa_densitymatrix_expecPauliString(rho.handle, coeffs.data(), (unsigned)coeffs.size(),
                                                 allPaulis.data(), out), "expecPauliString");
    return Amp(out[0], out[1]);
}

static DensityMatrix distributed_densitymatrix_partialTrace(DensityMatrix& inRho, NatArray targets) {       // :347-406
    fsa_state* out = nullptr;
    fsa_check(fsa_densitymatrix_partialTrace(inRho.handle, targets.data(), (unsigned)targets.size(), &out),
              "partialTrace");
    return DensityMatrix::adopt(out);
}

#endif  // DISTRIBUTED_DENSITYMATRIX_HPP
