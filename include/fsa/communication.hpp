// communication.hpp -- drop-in for the reference's src/communication.hpp at
// one rank on one GPU: rank 0 of 1, barriers synchronise the device, the
// amplitude reduction is the identity.  The point-to-point exchanges are only
// reached with more than one rank and are not part of this build.
#ifndef COMMUNICATION_HPP
#define COMMUNICATION_HPP

#include "types.hpp"
#include "fsa_check.hpp"

static inline void comm_init() { fsa_check(fsa_comm_init(), "comm_init"); }            // :16-21
static inline void comm_end() { fsa_check(fsa_comm_end(), "comm_end"); }               // :24-27
static inline Nat comm_getRank() { return fsa_comm_getRank(); }                        // :30-34
static inline Nat comm_getNumNodes() { return fsa_comm_getNumNodes(); }                // :37-41
static inline void comm_synch() { fsa_check(fsa_comm_synch(), "comm_synch"); }         // :44-47
static inline void comm_reduceAmp(Amp& amp) {                                          // :172-177
    double v[2] = {amp.real(), amp.imag()};
    fsa_check(fsa_comm_reduceAmp(v), "comm_reduceAmp");
    amp = Amp(v[0], v[1]);
}

#endif  // COMMUNICATION_HPP
