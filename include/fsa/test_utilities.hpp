// test_utilities.hpp -- drop-in for the reference's tests/test_utilities.hpp:
// the test harness's random generators, printers and dense checker, plus the
// StateVector / DensityMatrix members that states.hpp declares and the
// reference defines here (:419-524).  Catch2 is not included: a test file
// includes it itself.
//
// The random generators consume rand() in the reference's order, so one seed
// gives the same gates, targets and states under either library.  The dense
// checker builds each full 2^n x 2^n matrix directly from index bits.
//
// agreesWith compares |Re| and |Im| of each difference with tol; the
// reference compares the signed parts (:479, :516), so a negative difference
// never fails there.  A NaN difference does not fail in either.
#ifndef TEST_UTILITIES_HPP
#define TEST_UTILITIES_HPP

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "bit_maths.hpp"
#include "communication.hpp"
#include "misc.hpp"
#include "states.hpp"
#include "types.hpp"

static Real PI = 3.14159265358979323846;                                       // :19

static AmpMatrix matrI = {{1, 0}, {0, 1}};                                     // :21-24
static AmpMatrix matrX = {{0, 1}, {1, 0}};
static AmpMatrix matrY = {{0, Amp(0, -1)}, {Amp(0, 1), 0}};
static AmpMatrix matrZ = {{1, 0}, {0, -1}};

// ---- printing (:27-80) -----------------------------------------------------

static inline void printMatrix(AmpMatrix matrix) {
    for (const AmpArray& row : matrix) {
        for (const Amp& v : row) std::cout << v << " ";
        std::cout << "\n";
    }
}

template <typename T>
static inline void printArray(std::vector<T> array) {
    std::cout << "{";
    for (size_t i = 0; i < array.size(); i++) {
        if (i) std::cout << " ";
        std::cout << array[i];
    }
    std::cout << "}\n" << std::flush;
}

static inline void rootNodePrint(std::string msg) {
    comm_synch();
    std::cout << std::flush;
    if (comm_getRank() == 0) std::cout << msg << "\n" << std::flush;
    comm_synch();
}

template <typename T>
static inline void allNodesPrintLocalArray(std::vector<T> array) {
    for (Nat r = 0; r < comm_getNumNodes(); r++) {
        if (r == comm_getRank()) {
            std::cout << "  rank: " << r << "\n    ";
            printArray(array);
        }
        comm_synch();
    }
    std::cout << std::flush;
}

template <typename T>
static inline void rootNodePrintLocalArray(std::vector<T> array) {
    comm_synch();
    if (comm_getRank() == 0) printArray(array);
    std::cout << std::flush;
    comm_synch();
}

// ---- random inputs (:83-221) -------------------------------------------------

static inline Nat getRandomNat(Nat minInclusive, Nat maxExclusive) {
    assert(maxExclusive > minInclusive);
    return minInclusive + (Nat)(rand() % (maxExclusive - minInclusive));
}

static inline Real getRandomReal(Real minInclusive, Real maxExclusive) {
    const Real u = rand() / (Real)RAND_MAX;
    return minInclusive + u * (maxExclusive - minInclusive);
}

static inline RealArray getRandomRealArray(Real minInclusive, Real maxExclusive, Nat numElems) {
    assert(maxExclusive > minInclusive && numElems > 0);
    RealArray out(numElems);
    for (Real& x : out) x = getRandomReal(minInclusive, maxExclusive);
    return out;
}

// one complex normal sample by the Box-Muller transform of two rand() draws;
// the angle uses the reference's 8-digit pi (:110) so that the samples match
static inline Amp getRandomAmp() {
    const Real u1 = rand() / (Real)RAND_MAX;
    const Real u2 = rand() / (Real)RAND_MAX;
    const Real radius = std::sqrt(-2 * std::log(u1));
    const Real angle = 2 * 3.14159265 * u2;
    return Amp(radius * std::cos(angle), radius * std::sin(angle));
}

static inline AmpArray getRandomArray(Index dim) {
    assert(dim > 0);
    AmpArray out(dim);
    for (Amp& a : out) a = getRandomAmp();
    return out;
}

static inline AmpMatrix getRandomMatrix(Index dim) {
    assert(dim > 0);
    AmpMatrix m = getZeroMatrix(dim);
    for (AmpArray& row : m)
        for (Amp& a : row) a = getRandomAmp();
    return m;
}

static inline MatrixArray getRandomMatrices(Index dim, Nat numMatrices) {
    assert(dim > 0 && numMatrices > 0);
    MatrixArray out(numMatrices);
    for (AmpMatrix& m : out) m = getRandomMatrix(dim);
    return out;
}

static inline NatArray getRandomNatArray(Nat minIncl, Nat maxExcl, Nat numElem) {
    assert(maxExcl > minIncl && numElem > 0);
    NatArray out(numElem);
    for (Nat& x : out) x = getRandomNat(minIncl, maxExcl);
    return out;
}

// numElem distinct values of [minIncl, maxExcl): the range shuffled by
// 10 x (range size) random transpositions, first numElem kept
static inline NatArray getRandomUniqueNatArray(Nat minIncl, Nat maxExcl, Nat numElem) {
    assert(maxExcl > minIncl && numElem > 0);
    const Nat n = maxExcl - minIncl;
    assert(numElem <= n);
    NatArray all(n);
    for (Nat i = 0; i < n; i++) all[i] = minIncl + i;
    for (Nat rep = 0; rep < 10 * n; rep++) {
        const Nat a = getRandomNat(0, n);
        const Nat b = getRandomNat(0, n);
        std::swap(all[a], all[b]);
    }
    return NatArray(all.begin(), all.begin() + numElem);
}

// as above, with `exclude` (if drawn) replaced by the smallest value of the
// range that was not drawn
static inline NatArray getRandomUniqueNatArray(Nat minIncl, Nat maxExcl, Nat numElem, Nat exclude) {
    assert(exclude >= minIncl && exclude < maxExcl);
    assert(maxExcl - minIncl > numElem);
    NatArray out = getRandomUniqueNatArray(minIncl, maxExcl, numElem);
    auto hit = std::find(out.begin(), out.end(), exclude);
    if (hit != out.end()) {
        Nat v = minIncl;
        while (std::find(out.begin(), out.end(), v) != out.end()) v++;
        *hit = v;
    }
    return out;
}

static inline void ensureNotAllPauliZ(NatArray& paulis) {
    if (std::all_of(paulis.begin(), paulis.end(), [](Nat p) { return p == 3; })) paulis[0] = getRandomNat(1, 3);
}

// ---- dense matrices (:224-413) -----------------------------------------------

// exp(i angle P) = cos(angle) 1 + i sin(angle) P for a Pauli tensor P (P^2 = 1)
static inline AmpMatrix getExponentialOfPauliTensor(Real angle, AmpMatrix pauliTensor) {
    AmpMatrix e = Amp(0, std::sin(angle)) * pauliTensor;
    for (Index d = 0; d < e.size(); d++) e[d][d] += std::cos(angle);
    return e;
}

static inline void setSubMatrix(AmpMatrix& dest, AmpMatrix sub, Index r, Index c) {
    assert(r + sub.size() <= dest.size() && c + sub.size() <= dest.size());
    for (Index i = 0; i < sub.size(); i++)
        for (Index j = 0; j < sub.size(); j++) dest[r + i][c + j] = sub[i][j];
}

// the permutation matrix exchanging bits qb1 and qb2 of an numQb-bit index
static inline AmpMatrix getFullSwapMatrix(Nat qb1, Nat qb2, Nat numQb) {
    assert(qb1 < numQb && qb2 < numQb);
    const Index dim = powerOf2(numQb);
    AmpMatrix m = getZeroMatrix(dim);
    for (Index c = 0; c < dim; c++) {
        Index r = c;
        if (getBit(c, qb1) != getBit(c, qb2)) r = flipBit(flipBit(c, qb1), qb2);
        m[r][c] = 1;
    }
    return m;
}

static inline bool containsDuplicate(NatArray list1, NatArray list2) {
    NatArray all = list1;
    all.insert(all.end(), list2.begin(), list2.end());
    std::sort(all.begin(), all.end());
    return std::adjacent_find(all.begin(), all.end()) != all.end();
}

// the numQb-qubit matrix of gateMatr on targs (bit i of the gate's index is
// qubit targs[i]) controlled on every qubit of ctrls being 1
static inline AmpMatrix getFullGateMatrix(AmpMatrix gateMatr, NatArray ctrls, NatArray targs, Nat numQb) {
    assert(!targs.empty());
    assert(ctrls.size() + targs.size() <= numQb);
    assert(gateMatr.size() == powerOf2((Nat)targs.size()));
    assert(!containsDuplicate(ctrls, targs));
    for (Nat q : targs) assert(q < numQb);
    for (Nat q : ctrls) assert(q < numQb);
    const Index dim = powerOf2(numQb);
    const Index tmask = getBitMask(targs), cmask = getBitMask(ctrls);
    auto gather = [&](Index x) {
        Index g = 0;
        for (Nat i = 0; i < targs.size(); i++) g |= (Index)getBit(x, targs[i]) << i;
        return g;
    };
    AmpMatrix m = getZeroMatrix(dim);
    for (Index r = 0; r < dim; r++)
        for (Index c = 0; c < dim; c++) {
            if ((r & ~tmask) != (c & ~tmask)) continue;
            if ((c & cmask) != cmask) m[r][c] = (r == c) ? 1 : 0;
            else m[r][c] = gateMatr[gather(r)][gather(c)];
        }
    return m;
}

static inline void applyGateToLocalState(AmpArray& state, NatArray ctrls, NatArray targs, AmpMatrix gateMatr) {
    const AmpMatrix full = getFullGateMatrix(gateMatr, ctrls, targs, logBase2(state.size()));
    state = full * state;
}

static inline void applyGateToLocalState(AmpMatrix& state, NatArray ctrls, NatArray targs, AmpMatrix gateMatr) {
    const AmpMatrix full = getFullGateMatrix(gateMatr, ctrls, targs, logBase2(state.size()));
    state = full * state * getDaggerMatrix(full);
}

// sum_K K rho K^dagger
static inline void applyKrausMapToLocalState(AmpMatrix& state, NatArray targets, MatrixArray krausOps) {
    const AmpMatrix rho = state;
    state = getZeroMatrix(state.size());
    for (const AmpMatrix& K : krausOps) {
        AmpMatrix term = rho;
        applyGateToLocalState(term, {}, targets, K);
        state = state + term;
    }
}

// b on the low (fast) index: prod[r + |b| i][c + |b| j] = a[i][j] b[r][c]
static inline AmpMatrix getKroneckerProduct(AmpMatrix a, AmpMatrix b) { return a % b; }

// Pauli paulis[q] on qubit q (bit q of the index)
static inline AmpMatrix getKroneckerProductOfPaulis(NatArray paulis) {
    const AmpMatrix* m[4] = {&matrI, &matrX, &matrY, &matrZ};
    AmpMatrix prod = {{1}};
    for (Nat p : paulis) {
        assert(p <= 3);
        prod = getKroneckerProduct(*m[p], prod);
    }
    return prod;
}

// ---- the state members states.hpp declares (:419-524) ----------------------

// fresh unnormalised normal amplitudes; at one rank the local amplitudes are all
// of them, drawn in the reference's order and uploaded to the device
inline void StateVector::setRandomAmps() {
    comm_synch();
    amps = getRandomArray(numAmpsPerNode);
    syncToDevice();
}

inline void StateVector::printAmps() {
    comm_synch();
    syncFromDevice();
    std::printf("rank %u:\n", rank);
    std::fflush(stdout);
    printArray(amps);
    comm_synch();
}

inline bool StateVector::agreesWith(AmpArray ref, Real tol) {
    assert(ref.size() == powerOf2(numQubits));
    const AmpArray all = getAllVecAmps();
    for (Index i = 0; i < ref.size(); i++) {
        const Amp d = ref[i] - all[i];
        if (std::abs(d.real()) > tol || std::abs(d.imag()) > tol) {
            if (rank == 0) std::printf("disagreement of (%g) + i(%g)\n", d.real(), d.imag());
            return false;
        }
    }
    return true;
}

inline bool DensityMatrix::agreesWith(AmpMatrix ref, Real tol) {
    assert(ref.size() == powerOf2(numQubits));
    const AmpMatrix all = getAllMatrAmps();
    for (Index r = 0; r < ref.size(); r++)
        for (Index c = 0; c < ref.size(); c++) {
            const Amp d = ref[r][c] - all[r][c];
            if (std::abs(d.real()) > tol || std::abs(d.imag()) > tol) {
                if (rank == 0) std::printf("disagreement of (%g) + i(%g)\n", d.real(), d.imag());
                return false;
            }
        }
    return true;
}

#endif  // TEST_UTILITIES_HPP
