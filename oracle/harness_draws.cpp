// harness_draws.cpp -- TEST INFRASTRUCTURE.  harness_draws(path) writes the
// test harness's random draws and dense-checker matrices for fixed seeds to
// `path`, one number per line.  Built twice as a shared library: against the
// reference's tests/test_utilities.hpp (oracle/Makefile, target
// _ref/libharness_ref.so; it needs MPI for the state members it defines) and
// against the drop-in include/fsa/test_utilities.hpp.  The two outputs must be
// equal, so that one seed gives the same trials under either harness.
#include <cstdio>

#include "types.hpp"
#include "states.hpp"
#include "test_utilities.hpp"

static FILE* g_out = nullptr;
static void put(Real x) { std::fprintf(g_out, "%.17g\n", x); }
static void put(const Amp& a) { put(a.real()); put(a.imag()); }
static void put(const AmpMatrix& m) {
    for (const AmpArray& row : m)
        for (const Amp& a : row) put(a);
}
static void put(const NatArray& v) {
    for (Nat x : v) std::fprintf(g_out, "%u\n", x);
}

extern "C" int harness_draws(const char* path) {
    g_out = std::fopen(path, "w");
    if (!g_out) return 1;
    for (unsigned seed : {1u, 7u, 20261203u}) {
        srand(seed);
        std::fprintf(g_out, "%u\n", getRandomNat(0, 6));
        put(getRandomReal(-10, 10));
        put(getRandomAmp());
        for (Real x : getRandomRealArray(-1, 1, 5)) put(x);
        for (const Amp& a : getRandomArray(8)) put(a);
        put(getRandomMatrix(4));
        for (const AmpMatrix& m : getRandomMatrices(2, 3)) put(m);
        put(getRandomNatArray(0, 4, 10));
        put(getRandomUniqueNatArray(0, 6, 3));
        put(getRandomUniqueNatArray(0, 9, 5, 2));
        put(getRandomUniqueNatArray(3, 8, 4, 3));
        NatArray zs = {3, 3, 3};
        ensureNotAllPauliZ(zs);
        put(zs);
    }
    srand(5);
    AmpMatrix g = getRandomMatrix(4);
    put(getFullGateMatrix(g, {1}, {3, 0}, 4));
    put(getFullGateMatrix(getRandomMatrix(2), {0, 2}, {1}, 3));
    put(getFullSwapMatrix(0, 3, 4));
    put(getFullSwapMatrix(2, 1, 3));
    put(getKroneckerProductOfPaulis({1, 2, 3}));
    put(getExponentialOfPauliTensor(0.3, getKroneckerProductOfPaulis({2, 0, 1})));
    put(getKroneckerProduct(getRandomMatrix(2), getRandomMatrix(4)));
    AmpArray v = getRandomArray(16);
    applyGateToLocalState(v, {2}, {0, 3}, g);
    for (const Amp& a : v) put(a);
    AmpMatrix rho = getRandomMatrix(8);
    applyKrausMapToLocalState(rho, {2, 0}, getRandomMatrices(4, 2));
    put(rho);
    return std::fclose(g_out) == 0 ? 0 : 1;
}
