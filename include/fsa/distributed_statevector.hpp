// distributed_statevector.hpp -- drop-in for the reference's
// src/distributed_statevector.hpp: same names and signatures, bodies call the
// libfsa C ABI (one rank, one GPU; every op takes the local path).
#ifndef DISTRIBUTED_STATEVECTOR_HPP
#define DISTRIBUTED_STATEVECTOR_HPP

#include "local_statevector.hpp"
#include "misc.hpp"

static void distributed_statevector_oneTargGate(StateVector& psi, Nat target, AmpMatrix gate) {      // :18-40
    fsa_check(fsa_statevector_oneTargGate(psi.handle, target, fsa_dropin::flat(gate).data()), "oneTargGate");
}

static void distributed_statevector_manyCtrlOneTargGate(StateVector& psi, NatArray controls, Nat target,
                                                        AmpMatrix gate) {                             // :81-106
    fsa_check(fsa_statevector_manyCtrlOneTargGate(psi.handle, controls.data(), (unsigned)controls.size(), target,
                                                  fsa_dropin::flat(gate).data()), "manyCtrlOneTargGate");
}

static void distributed_statevector_swapGate(StateVector& psi, Nat qb1, Nat qb2) {                     // :109-187
    fsa_check(fsa_statevector_swapGate(psi.handle, qb1, qb2), "swapGate");
}

static void distributed_statevector_manyTargGate(StateVector& psi, NatArray targets, AmpMatrix gate) { // :190-224
    fsa_check(fsa_statevector_manyTargGate(psi.handle, targets.data(), (unsigned)targets.size(),
                                           fsa_dropin::flat(gate).data()), "manyTargGate");
}

static void distributed_statevector_pauliTensor(StateVector& psi, NatArray targets, NatArray paulis) { // :279-284
    assert(targets.size() == paulis.size());
    fsa_check(fsa_statevector_pauliTensor(psi.handle, targets.data(), paulis.data(), (unsigned)targets.size()),
              "pauliTensor");
}

static void distributed_statevector_pauliGadget(StateVector& psi, NatArray targets, NatArray paulis,
                                                Real theta) {                                        // :287-292
    assert(targets.size() == paulis.size());
    fsa_check(fsa_statevector_pauliGadget(psi.handle, targets.data(), paulis.data(), (unsigned)targets.size(),
                                          theta), "pauliGadget");
}

static void distributed_statevector_phaseGadget(StateVector& psi, NatArray targets, Real theta) {     // :295-298
    fsa_check(fsa_statevector_phaseGadget(psi.handle, targets.data(), (unsigned)targets.size(), theta),
              "phaseGadget");
}

#endif  // DISTRIBUTED_STATEVECTOR_HPP
