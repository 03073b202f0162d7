// local_statevector.hpp -- drop-in for the reference's src/local_statevector.hpp:
// same names and signatures.  At one rank a process's local amplitudes are
// the whole state, so each local routine is the libfsa operation on the
// state's device amplitudes (its host mirror `amps` is refreshed by
// getAllVecAmps()).
#ifndef LOCAL_STATEVECTOR_HPP
#define LOCAL_STATEVECTOR_HPP

#include <vector>

#include "bit_maths.hpp"
#include "fsa_check.hpp"
#include "states.hpp"
#include "types.hpp"

namespace fsa_dropin {
// row-major interleaved (re, im) copy of a square AmpMatrix, the layout of
// every gate argument of the C ABI
inline std::vector<double> flat(const AmpMatrix& m) {
    std::vector<double> f;
    f.reserve(2 * m.size() * m.size());
    for (const AmpArray& row : m)
        for (const Amp& v : row) { f.push_back(v.real()); f.push_back(v.imag()); }
    return f;
}
inline void one_rank(const StateVector& psi) {
    // the multi-rank exchanges are not part of this build
    assert(psi.numNodes == 1);
    (void)psi;
}
}  // namespace fsa_dropin

static void local_statevector_oneTargGate(StateVector& psi, Nat target, AmpMatrix gate) {                  // :14-29
    fsa_dropin::one_rank(psi);
    fsa_check(fsa_statevector_oneTargGate(psi.handle, target, fsa_dropin::flat(gate).data()), "oneTargGate");
}

static void local_statevector_manyCtrlOneTargGate(StateVector& psi, NatArray controls, Nat target,
                                                  AmpMatrix gate) {                                        // :32-51
    fsa_dropin::one_rank(psi);
    fsa_check(fsa_statevector_manyCtrlOneTargGate(psi.handle, controls.data(), (unsigned)controls.size(), target,
                                                  fsa_dropin::flat(gate).data()), "manyCtrlOneTargGate");
}

static void local_statevector_swapGate(StateVector& psi, Nat qb1, Nat qb2) {                                // :54-69
    fsa_dropin::one_rank(psi);
    fsa_check(fsa_statevector_swapGate(psi.handle, qb1, qb2), "swapGate");
}

static void local_statevector_manyTargGate(StateVector& psi, NatArray targets, AmpMatrix gate) {            // :72-99
    fsa_dropin::one_rank(psi);
    fsa_check(fsa_statevector_manyTargGate(psi.handle, targets.data(), (unsigned)targets.size(),
                                           fsa_dropin::flat(gate).data()), "manyTargGate");
}

// :102-135.  targs (the X/Y targets, whose highest bit maskXY holds) only
// orders the reference's loops; the operation is fixed by the masks:
// psi <- thisAmpFac psi + otherAmpFac powI X^maskXY Z^maskYZ psi.
static void local_statevector_pauliTensorOrGadget_subroutine(StateVector& psi, NatArray targs, Amp powI, Index maskXY,
                                                             Index maskYZ, Amp thisAmpFac, Amp otherAmpFac) {
    fsa_dropin::one_rank(psi);
    (void)targs;
    const double p[2] = {powI.real(), powI.imag()};
    const double t[2] = {thisAmpFac.real(), thisAmpFac.imag()};
    const double o[2] = {otherAmpFac.real(), otherAmpFac.imag()};
    fsa_check(fsa_local_statevector_pauliTensorOrGadget(psi.handle, maskXY, maskYZ, p, t, o),
              "pauliTensorOrGadget_subroutine");
}

static void local_statevector_phaseGadget(StateVector& psi, NatArray targets, Real theta) {                 // :138-153
    fsa_dropin::one_rank(psi);
    fsa_check(fsa_statevector_phaseGadget(psi.handle, targets.data(), (unsigned)targets.size(), theta),
              "phaseGadget");
}

#endif  // LOCAL_STATEVECTOR_HPP
