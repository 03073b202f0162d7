// test_harness.cpp -- the drop-in test harness (include/fsa/test_utilities.hpp)
// checked on the CPU: its dense checker against the oracle's operations, and
// its random generators against their contracts.  Needs no GPU (libfsa is
// linked for the state members but never called).
// Exit status 0 iff every check passes.
#include <cmath>
#include <cstdio>
#include <set>

#include "local_statevector.hpp"
#include "test_utilities.hpp"
#include "../../oracle/fsa_oracle.h"

static int g_failures = 0;
static void check(const char* name, bool ok) {
    if (!ok) {
        if (g_failures < 20) std::printf("FAIL %s\n", name);
        g_failures++;
    }
}

static double relErr(const AmpArray& a, const AmpArray& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) { num += std::norm(a[i] - b[i]); den += std::norm(b[i]); }
    return std::sqrt(num / den);
}
static std::vector<double> flat(const AmpMatrix& m) { return fsa_dropin::flat(m); }
static double* raw(AmpArray& v) { return reinterpret_cast<double*>(v.data()); }

// the density matrix's flat vector: rho[r][c] = v[r + 2^N c]
static AmpArray vecOf(const AmpMatrix& m) {
    const Index d = m.size();
    AmpArray v(d * d);
    for (Index r = 0; r < d; r++)
        for (Index c = 0; c < d; c++) v[r + d * c] = m[r][c];
    return v;
}

int main() {
    srand(20261203);
    const Nat n = 6, N = 4;
    for (int t = 0; t < 200; t++) {
        // controlled one-target gate
        {
            AmpArray psi = getRandomArray(powerOf2(n)), ref = psi;
            Nat numCtrls = getRandomNat(0, n - 1);
            Nat targ = getRandomNat(0, n);
            NatArray ctrls = numCtrls ? getRandomUniqueNatArray(0, n, numCtrls, targ) : NatArray{};
            AmpMatrix g = getRandomMatrix(2);
            applyGateToLocalState(psi, ctrls, {targ}, g);
            orc_sv_manyCtrlOneTargGate(raw(ref), n, ctrls.data(), numCtrls, targ, flat(g).data());
            check("applyGateToLocalState(controlled) vs oracle", relErr(psi, ref) <= 1e-13);
        }
        // many-target gate
        {
            AmpArray psi = getRandomArray(powerOf2(n)), ref = psi;
            Nat k = getRandomNat(1, n + 1);
            NatArray targs = getRandomUniqueNatArray(0, n, k);
            AmpMatrix g = getRandomMatrix(powerOf2(k));
            applyGateToLocalState(psi, {}, targs, g);
            orc_sv_manyTargGate(raw(ref), n, targs.data(), k, flat(g).data());
            check("applyGateToLocalState(many targets) vs oracle", relErr(psi, ref) <= 1e-13);
        }
        // swap
        {
            AmpArray psi = getRandomArray(powerOf2(n)), ref = psi;
            NatArray q = getRandomUniqueNatArray(0, n, 2);
            psi = getFullSwapMatrix(q[0], q[1], n) * psi;
            orc_sv_swapGate(raw(ref), n, q[0], q[1]);
            check("getFullSwapMatrix vs oracle", relErr(psi, ref) == 0);
        }
        // exp(i theta P) on every qubit in order
        {
            AmpArray psi = getRandomArray(powerOf2(n)), ref = psi;
            NatArray paulis = getRandomNatArray(1, 4, n), all(n);
            for (Nat q = 0; q < n; q++) all[q] = q;
            Real th = getRandomReal(-PI, PI);
            psi = getExponentialOfPauliTensor(th, getKroneckerProductOfPaulis(paulis)) * psi;
            orc_sv_pauliGadget(raw(ref), n, all.data(), paulis.data(), n, th);
            check("getExponentialOfPauliTensor x getKroneckerProductOfPaulis vs oracle", relErr(psi, ref) <= 1e-13);
        }
        // density matrix: U rho U^dagger and a Kraus map
        {
            AmpMatrix rho = getRandomMatrix(powerOf2(N));
            AmpArray ref = vecOf(rho);
            Nat k = getRandomNat(1, 3);
            NatArray targs = getRandomUniqueNatArray(0, N, k);
            AmpMatrix g = getRandomMatrix(powerOf2(k));
            applyGateToLocalState(rho, {}, targs, g);
            orc_dm_manyTargGate(raw(ref), N, targs.data(), k, flat(g).data());
            check("applyGateToLocalState(density matrix) vs oracle", relErr(vecOf(rho), ref) <= 1e-13);

            MatrixArray ops = getRandomMatrices(powerOf2(k), getRandomNat(1, 4));
            std::vector<double> all;
            for (auto& K : ops) { auto f = flat(K); all.insert(all.end(), f.begin(), f.end()); }
            ref = vecOf(rho);
            applyKrausMapToLocalState(rho, targs, ops);
            orc_dm_krausMap(raw(ref), N, all.data(), (unsigned)ops.size(), targs.data(), k);
            check("applyKrausMapToLocalState vs oracle", relErr(vecOf(rho), ref) <= 1e-13);
        }
        // Tr(P rho) with the Kronecker product of Paulis
        {
            AmpMatrix rho = getRandomMatrix(powerOf2(N));
            AmpArray v = vecOf(rho);
            NatArray paulis = getRandomNatArray(0, 4, N);
            AmpMatrix pr = getKroneckerProductOfPaulis(paulis) * rho;
            Amp tr = 0;
            for (Index d = 0; d < pr.size(); d++) tr += pr[d][d];
            double coeff = 1, e[2];
            orc_dm_expecPauliString(raw(v), N, &coeff, 1, paulis.data(), e);
            check("getKroneckerProductOfPaulis trace vs oracle", std::abs(tr - Amp(e[0], e[1])) <= 1e-12 * (1 + std::abs(tr)));
        }
        // generators' contracts
        {
            Nat lo = getRandomNat(0, 5), hi = lo + getRandomNat(2, 9);
            Nat num = getRandomNat(1, hi - lo);
            Nat excl = getRandomNat(lo, hi);
            NatArray u = getRandomUniqueNatArray(lo, hi, num, excl);
            std::set<Nat> s(u.begin(), u.end());
            bool inRange = std::all_of(u.begin(), u.end(), [&](Nat x) { return x >= lo && x < hi; });
            check("getRandomUniqueNatArray: distinct, in range, without the excluded value",
                  s.size() == u.size() && inRange && !s.count(excl));
            Real r = getRandomReal(-3, 5);
            check("getRandomReal in range", r >= -3 && r <= 5);
            NatArray z(3, 3);
            ensureNotAllPauliZ(z);
            check("ensureNotAllPauliZ", z[0] == 1 || z[0] == 2);
        }
    }
    // the controlled gate's full matrix on a hand-checked case: CNOT, control 1, target 0
    {
        AmpMatrix m = getFullGateMatrix(matrX, {1}, {0}, 2);
        AmpMatrix cnot = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}};
        check("getFullGateMatrix(CNOT)", m == cnot);
        check("containsDuplicate", containsDuplicate({1, 2}, {3, 2}) && !containsDuplicate({1}, {0, 2}));
    }
    std::printf("harness checks: %d failures\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
