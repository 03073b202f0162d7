// misc.hpp -- drop-in for the host helpers of the reference's src/misc.hpp.
#ifndef MISC_HPP
#define MISC_HPP

#include "bit_maths.hpp"
#include "types.hpp"

// misc.hpp:16-38: element of the Pauli tensor at a flat density-matrix index
static inline Amp getPauliTensorElem(const Nat* pauliCodes, Nat numQubits, Index flatInd) {
    static const Amp m[4][2][2] = {{{1, 0}, {0, 1}}, {{0, 1}, {1, 0}},
                                   {{0, Amp(0, -1)}, {Amp(0, 1), 0}}, {{1, 0}, {0, -1}}};
    Amp e = 1.;
    for (Nat q = 0; q < numQubits; q++) e *= m[pauliCodes[q]][getBit(flatInd, q + numQubits)][getBit(flatInd, q)];
    return e;
}

// misc.hpp:47-55
static inline bool containsOddNumY(const NatArray& paulis) {
    bool odd = false;
    for (Nat p : paulis) if (p == Y) odd = !odd;
    return odd;
}

// misc.hpp:58-81: sum_K conj(K) (x) K with row i*d + k, column j*d + l
static inline AmpMatrix getSuperoperator(const MatrixArray& krausOps) {
    Index d = krausOps[0].size(), D = d * d;
    AmpMatrix s = getZeroMatrix(D);
    for (const AmpMatrix& K : krausOps)
        for (Index i = 0; i < d; i++)
            for (Index j = 0; j < d; j++)
                for (Index k = 0; k < d; k++)
                    for (Index l = 0; l < d; l++) s[i * d + k][j * d + l] += std::conj(K[i][j]) * K[k][l];
    return s;
}

#endif  // MISC_HPP
