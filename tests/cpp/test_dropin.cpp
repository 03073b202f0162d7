// test_dropin.cpp -- the reference's test style (tests/tests_statevector.hpp,
// tests/tests_densitymatrix.hpp) run through the drop-in headers: each trial
// prepares a random state, applies one operation with the reference's
// function names (executed by libfsa on the GPU) and with the CPU oracle,
// and compares.  Built with plain g++ (the boundary is a C ABI); needs a GPU.
// Exit status 0 iff every trial agrees to 1e-12 relative 2-norm.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>

#include "distributed_densitymatrix.hpp"
#include "local_densitymatrix.hpp"
#include "local_statevector.hpp"
#include "test_utilities.hpp"
#include "../../oracle/fsa_oracle.h"

static const int TRIALS = 200;
static const Nat NQ_SV = 6, NQ_DM = 5;          // tests_statevector.hpp:13, tests_densitymatrix.hpp:151
static int g_failures = 0;

static double urand() { return rand() / (double)RAND_MAX; }
static Amp grand() {
    double a = urand(), b = urand();
    if (a < 1e-300) a = 1e-300;
    return Amp(std::sqrt(-2 * std::log(a)) * std::cos(6.283185307179586 * b),
               std::sqrt(-2 * std::log(a)) * std::sin(6.283185307179586 * b));
}
static AmpMatrix randMatrix(Index d) {
    AmpMatrix m = getZeroMatrix(d);
    for (auto& r : m) for (auto& v : r) v = grand();
    return m;
}
static NatArray uniqueNats(Nat n, Nat k) {
    NatArray all(n);
    for (Nat i = 0; i < n; i++) all[i] = i;
    for (Nat i = 0; i < n; i++) std::swap(all[i], all[i + rand() % (n - i)]);
    return NatArray(all.begin(), all.begin() + k);
}
static std::vector<double> flatm(const AmpMatrix& m) { return fsa_dropin::flat(m); }

static double relErr(const AmpArray& a, const AmpArray& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) { num += std::norm(a[i] - b[i]); den += std::norm(b[i]); }
    return std::sqrt(num / den);
}

// amplitude idx -> the index whose bit perm[k] is bit k of idx
static AmpArray permuteBits(const AmpArray& v, const std::vector<Nat>& perm) {
    AmpArray out(v.size());
    for (Index i = 0; i < v.size(); i++) {
        Index j = 0;
        for (Nat k = 0; k < perm.size(); k++) j |= (Index)getBit(i, k) << perm[k];
        out[j] = v[i];
    }
    return out;
}

static void check(const char* name, bool ok) {
    if (!ok) {
        if (g_failures < 20) std::printf("FAIL %s\n", name);
        g_failures++;
    }
}

template <class S>
static void trial(const char* name, S& state, const std::function<void(S&)>& gpu,
                  const std::function<void(double*)>& oracle) {
    for (auto& v : state.amps) v = grand();
    state.syncToDevice();
    AmpArray ref = state.amps;
    gpu(state);
    oracle(reinterpret_cast<double*>(ref.data()));
    double e = relErr(state.getAllVecAmps(), ref);
    if (!(e <= 1e-12)) {
        if (g_failures < 20) std::printf("FAIL %s: relative error %g\n", name, e);
        g_failures++;
    }
}

int main() {
    comm_init();
    srand(20261202);
    StateVector psi(NQ_SV);
    DensityMatrix rho(NQ_DM);
    for (int t = 0; t < TRIALS; t++) {
        // ---- statevector (tests_statevector.hpp:24-134)
        {
            AmpMatrix g = randMatrix(2);
            Nat targ = rand() % NQ_SV;
            trial<StateVector>("statevector_oneTargGate", psi,
                [&](StateVector& s) { distributed_statevector_oneTargGate(s, targ, g); },
                [&](double* r) { orc_sv_oneTargGate(r, NQ_SV, targ, flatm(g).data()); });
        }
        {
            AmpMatrix g = randMatrix(2);
            Nat nc = 1 + rand() % (NQ_SV - 1);
            NatArray q = uniqueNats(NQ_SV, nc + 1);
            Nat targ = q[0];
            NatArray ctrls(q.begin() + 1, q.end());
            trial<StateVector>("statevector_manyCtrlOneTargGate", psi,
                [&](StateVector& s) { distributed_statevector_manyCtrlOneTargGate(s, ctrls, targ, g); },
                [&](double* r) { orc_sv_manyCtrlOneTargGate(r, NQ_SV, ctrls.data(), nc, targ, flatm(g).data()); });
        }
        {
            NatArray q = uniqueNats(NQ_SV, 2);
            trial<StateVector>("statevector_swapGate", psi,
                [&](StateVector& s) { distributed_statevector_swapGate(s, q[0], q[1]); },
                [&](double* r) { orc_sv_swapGate(r, NQ_SV, q[0], q[1]); });
        }
        {
            Nat k = 1 + rand() % NQ_SV;
            NatArray targs = uniqueNats(NQ_SV, k);
            AmpMatrix g = randMatrix(powerOf2(k));
            trial<StateVector>("statevector_manyTargGate", psi,
                [&](StateVector& s) { distributed_statevector_manyTargGate(s, targs, g); },
                [&](double* r) { orc_sv_manyTargGate(r, NQ_SV, targs.data(), k, flatm(g).data()); });
        }
        {
            Nat k = 1 + rand() % NQ_SV;
            NatArray targs = uniqueNats(NQ_SV, k), paulis(k);
            for (auto& p : paulis) p = 1 + rand() % 3;
            Real th = -3.14159 + 6.28318 * urand();
            trial<StateVector>("statevector_pauliTensor", psi,
                [&](StateVector& s) { distributed_statevector_pauliTensor(s, targs, paulis); },
                [&](double* r) { orc_sv_pauliTensor(r, NQ_SV, targs.data(), paulis.data(), k); });
            trial<StateVector>("statevector_pauliGadget", psi,
                [&](StateVector& s) { distributed_statevector_pauliGadget(s, targs, paulis, th); },
                [&](double* r) { orc_sv_pauliGadget(r, NQ_SV, targs.data(), paulis.data(), k, th); });
            trial<StateVector>("statevector_phaseGadget", psi,
                [&](StateVector& s) { distributed_statevector_phaseGadget(s, targs, th); },
                [&](double* r) { orc_sv_phaseGadget(r, NQ_SV, targs.data(), k, th); });
        }
        // ---- density matrix (tests_densitymatrix.hpp:26-282)
        {
            Nat k = 1 + rand() % 3;
            NatArray targs = uniqueNats(NQ_DM, k);
            AmpMatrix g = randMatrix(powerOf2(k));
            trial<DensityMatrix>("densitymatrix_manyTargGate", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_manyTargGate(s, targs, g); },
                [&](double* r) { orc_dm_manyTargGate(r, NQ_DM, targs.data(), k, flatm(g).data()); });
            MatrixArray ks;
            for (int i = 0; i < 1 + rand() % 4; i++) ks.push_back(randMatrix(powerOf2(k)));
            std::vector<double> kflat;
            for (auto& K : ks) { auto f = flatm(K); kflat.insert(kflat.end(), f.begin(), f.end()); }
            trial<DensityMatrix>("densitymatrix_krausMap", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_krausMap(s, ks, targs); },
                [&](double* r) { orc_dm_krausMap(r, NQ_DM, kflat.data(), (unsigned)ks.size(), targs.data(), k); });
        }
        {
            NatArray q = uniqueNats(NQ_DM, 2);
            Real p = 0.5 * urand();
            Nat k = 1 + rand() % NQ_DM;
            NatArray targs = uniqueNats(NQ_DM, k), paulis(k);
            for (auto& x : paulis) x = 1 + rand() % 3;
            Real th = -3.14159 + 6.28318 * urand();
            trial<DensityMatrix>("densitymatrix_swapGate", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_swapGate(s, q[0], q[1]); },
                [&](double* r) { orc_dm_swapGate(r, NQ_DM, q[0], q[1]); });
            trial<DensityMatrix>("densitymatrix_pauliTensor", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_pauliTensor(s, targs, paulis); },
                [&](double* r) { orc_dm_pauliTensor(r, NQ_DM, targs.data(), paulis.data(), k); });
            trial<DensityMatrix>("densitymatrix_pauliGadget", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_pauliGadget(s, targs, paulis, th); },
                [&](double* r) { orc_dm_pauliGadget(r, NQ_DM, targs.data(), paulis.data(), k, th); });
            trial<DensityMatrix>("densitymatrix_phaseGadget", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_phaseGadget(s, targs, th); },
                [&](double* r) { orc_dm_phaseGadget(r, NQ_DM, targs.data(), k, th); });
            trial<DensityMatrix>("densitymatrix_oneQubitDephasing", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_oneQubitDephasing(s, q[0], p); },
                [&](double* r) { orc_dm_oneQubitDephasing(r, NQ_DM, q[0], p); });
            trial<DensityMatrix>("densitymatrix_twoQubitDephasing", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_twoQubitDephasing(s, q[0], q[1], p); },
                [&](double* r) { orc_dm_twoQubitDephasing(r, NQ_DM, q[0], q[1], p); });
            trial<DensityMatrix>("densitymatrix_oneQubitDepolarising", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_oneQubitDepolarising(s, q[0], p); },
                [&](double* r) { orc_dm_oneQubitDepolarising(r, NQ_DM, q[0], p); });
            trial<DensityMatrix>("densitymatrix_twoQubitDepolarising", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_twoQubitDepolarising(s, q[0], q[1], p); },
                [&](double* r) { orc_dm_twoQubitDepolarising(r, NQ_DM, q[0], q[1], p); });
            trial<DensityMatrix>("densitymatrix_damping", rho,
                [&](DensityMatrix& s) { distributed_densitymatrix_damping(s, q[0], p); },
                [&](double* r) { orc_dm_damping(r, NQ_DM, q[0], p); });
        }
        {
            // tests_densitymatrix.hpp:354-380
            for (auto& v : rho.amps) v = grand();
            rho.syncToDevice();
            Nat T = 1 + rand() % 29;
            NatArray paulis(T * NQ_DM);
            for (auto& x : paulis) x = rand() % 4;
            RealArray coeffs(T);
            for (auto& c : coeffs) c = -10 + 20 * urand();
            Amp e1 = distributed_densitymatrix_expecPauliString(rho, coeffs, paulis);
            double e2[2];
            orc_dm_expecPauliString(reinterpret_cast<double*>(rho.amps.data()), NQ_DM, coeffs.data(), T, paulis.data(), e2);
            if (std::abs(e1 - Amp(e2[0], e2[1])) > 1e-12 * std::max(1.0, std::abs(Amp(e2[0], e2[1])))) {
                std::printf("FAIL densitymatrix_expecPauliString\n");
                g_failures++;
            }
            // tests_densitymatrix.hpp:383-418
            Nat k = 1 + rand() % NQ_DM;
            NatArray targs = uniqueNats(NQ_DM, k);
            DensityMatrix out = distributed_densitymatrix_partialTrace(rho, targs);
            AmpArray ref(out.numAmpsPerNode);
            orc_dm_partialTrace(reinterpret_cast<double*>(rho.amps.data()), NQ_DM, targs.data(), k,
                                reinterpret_cast<double*>(ref.data()));
            if (!(relErr(out.getAllVecAmps(), ref) <= 1e-12)) {
                std::printf("FAIL densitymatrix_partialTrace\n");
                g_failures++;
            }
        }
    }
    // ---- local_statevector.hpp / local_densitymatrix.hpp (one rank: the whole state)
    for (int t = 0; t < TRIALS; t++) {
        {
            AmpMatrix g = randMatrix(2);
            Nat targ = rand() % NQ_SV;
            trial<StateVector>("local_statevector_oneTargGate", psi,
                [&](StateVector& s) { local_statevector_oneTargGate(s, targ, g); },
                [&](double* r) { orc_sv_oneTargGate(r, NQ_SV, targ, flatm(g).data()); });
            NatArray q = uniqueNats(NQ_SV, 3);
            NatArray ctrls = {q[1], q[2]};
            trial<StateVector>("local_statevector_manyCtrlOneTargGate", psi,
                [&](StateVector& s) { local_statevector_manyCtrlOneTargGate(s, ctrls, q[0], g); },
                [&](double* r) { orc_sv_manyCtrlOneTargGate(r, NQ_SV, ctrls.data(), 2, q[0], flatm(g).data()); });
            trial<StateVector>("local_statevector_swapGate", psi,
                [&](StateVector& s) { local_statevector_swapGate(s, q[0], q[2]); },
                [&](double* r) { orc_sv_swapGate(r, NQ_SV, q[0], q[2]); });
            AmpMatrix g3 = randMatrix(8);
            trial<StateVector>("local_statevector_manyTargGate", psi,
                [&](StateVector& s) { local_statevector_manyTargGate(s, q, g3); },
                [&](double* r) { orc_sv_manyTargGate(r, NQ_SV, q.data(), 3, flatm(g3).data()); });
            Real th = -3.14159 + 6.28318 * urand();
            trial<StateVector>("local_statevector_phaseGadget", psi,
                [&](StateVector& s) { local_statevector_phaseGadget(s, q, th); },
                [&](double* r) { orc_sv_phaseGadget(r, NQ_SV, q.data(), 3, th); });
        }
        {
            // a psi + b P psi for the Pauli tensor P (powI = i^{#Y}, local_statevector.hpp:102-135),
            // by linearity from the oracle's pauliTensor
            Nat k = 1 + rand() % NQ_SV;
            NatArray targs = uniqueNats(NQ_SV, k), paulis(k);
            Index mxy = 0, myz = 0;
            Amp powI = 1;
            for (Nat i = 0; i < k; i++) {
                paulis[i] = 1 + rand() % 3;
                if (paulis[i] != 3) mxy |= powerOf2(targs[i]);
                if (paulis[i] != 1) myz |= powerOf2(targs[i]);
                if (paulis[i] == 2) powI *= Amp(0, 1);
            }
            Amp a = grand(), b = grand();
            trial<StateVector>("local_statevector_pauliTensorOrGadget_subroutine", psi,
                [&](StateVector& s) {
                    local_statevector_pauliTensorOrGadget_subroutine(s, targs, powI, mxy, myz, a, b);
                },
                [&](double* r) {
                    AmpArray in(reinterpret_cast<Amp*>(r), reinterpret_cast<Amp*>(r) + psi.numAmpsPerNode);
                    orc_sv_pauliTensor(r, NQ_SV, targs.data(), paulis.data(), k);
                    Amp* o = reinterpret_cast<Amp*>(r);
                    for (Index i = 0; i < in.size(); i++) o[i] = a * in[i] + b * o[i];
                });
        }
        {
            NatArray q = uniqueNats(NQ_DM, 2);
            Real p = 0.5 * urand();
            // the coefficients distributed_densitymatrix.hpp:106-108 passes
            trial<DensityMatrix>("local_densitymatrix_oneQubitDepolarising", rho,
                [&](DensityMatrix& s) {
                    local_densitymatrix_oneQubitDepolarising(s, q[0], 2 * p / 3, 1 - 2 * p / 3, 1 - 4 * p / 3);
                },
                [&](double* r) { orc_dm_oneQubitDepolarising(r, NQ_DM, q[0], p); });
            // :247-256: sorted qubits, shipped constants
            Nat lo = std::min(q[0], q[1]), hi = std::max(q[0], q[1]);
            trial<DensityMatrix>("local_densitymatrix_twoQubitDepolarising", rho,
                [&](DensityMatrix& s) {
                    local_densitymatrix_twoQubitDepolarising(s, lo, hi, lo + NQ_DM, hi + NQ_DM, 1 - 4 * p / 5,
                                                             4 * p / 15, -16 * p / 15);
                },
                [&](double* r) { orc_dm_twoQubitDepolarising_asShipped(r, NQ_DM, lo, hi, p); });
            // any four bits: relabel them onto (lo, hi, lo + N, hi + N), apply, relabel back
            NatArray bits = uniqueNats(2 * NQ_DM, 4);
            const Nat dest[4] = {lo, hi, lo + NQ_DM, hi + NQ_DM};
            std::vector<Nat> perm(2 * NQ_DM, 2 * NQ_DM), inv(2 * NQ_DM);
            std::vector<bool> used(2 * NQ_DM, false);
            for (int i = 0; i < 4; i++) { perm[bits[i]] = dest[i]; used[dest[i]] = true; }
            Nat nextFree = 0;
            for (Nat b2 = 0; b2 < 2 * NQ_DM; b2++)
                if (perm[b2] == 2 * NQ_DM) {
                    while (used[nextFree]) nextFree++;
                    perm[b2] = nextFree;
                    used[nextFree] = true;
                }
            for (Nat b2 = 0; b2 < 2 * NQ_DM; b2++) inv[perm[b2]] = b2;
            trial<DensityMatrix>("local_densitymatrix_twoQubitDepolarising(any bits)", rho,
                [&](DensityMatrix& s) {
                    local_densitymatrix_twoQubitDepolarising(s, bits[0], bits[1], bits[2], bits[3], 1 - 4 * p / 5,
                                                             4 * p / 15, -16 * p / 15);
                },
                [&](double* r) {
                    Amp* o = reinterpret_cast<Amp*>(r);
                    AmpArray v = permuteBits(AmpArray(o, o + rho.numAmpsPerNode), perm);
                    orc_dm_twoQubitDepolarising_asShipped(reinterpret_cast<double*>(v.data()), NQ_DM, lo, hi, p);
                    v = permuteBits(v, inv);
                    std::copy(v.begin(), v.end(), o);
                });
            trial<DensityMatrix>("local_densitymatrix_oneQubitDephasing", rho,
                [&](DensityMatrix& s) { local_densitymatrix_oneQubitDephasing(s, q[0], p); },
                [&](double* r) { orc_dm_oneQubitDephasing(r, NQ_DM, q[0], p); });
            trial<DensityMatrix>("local_densitymatrix_twoQubitDephasing", rho,
                [&](DensityMatrix& s) { local_densitymatrix_twoQubitDephasing(s, q[0], q[1], p); },
                [&](double* r) { orc_dm_twoQubitDephasing(r, NQ_DM, q[0], q[1], p); });
            trial<DensityMatrix>("local_densitymatrix_damping", rho,
                [&](DensityMatrix& s) { local_densitymatrix_damping(s, q[0], p); },
                [&](double* r) { orc_dm_damping(r, NQ_DM, q[0], p); });
            // :134-164 with the pairing distributed_densitymatrix.hpp:365-369 uses
            for (auto& v : rho.amps) v = grand();
            rho.syncToDevice();
            Nat k = 1 + rand() % NQ_DM;
            NatArray targs = uniqueNats(NQ_DM, k), pairs = targs;
            for (Nat& x : pairs) x += NQ_DM;
            DensityMatrix out = local_densitymatrix_partialTrace(rho, targs, pairs);
            AmpArray ref(out.numAmpsPerNode);
            orc_dm_partialTrace(reinterpret_cast<double*>(rho.amps.data()), NQ_DM, targs.data(), k,
                                reinterpret_cast<double*>(ref.data()));
            check("local_densitymatrix_partialTrace", relErr(out.getAllVecAmps(), ref) <= 1e-12);
        }
    }

    // ---- the test-utility members (test_utilities.hpp:437-524)
    {
        srand(7);
        psi.setRandomAmps();
        srand(7);
        AmpArray drawn = getRandomArray(psi.numAmpsPerNode);
        AmpArray held = psi.getAllVecAmps();
        check("setRandomAmps draws getRandomArray's amplitudes", held == drawn);
        check("agreesWith itself", psi.agreesWith(drawn));
        AmpArray off = drawn;
        off[3] -= Amp(2e-5, 0);
        check("agreesWith flags a negative difference", !psi.agreesWith(off));
        off = drawn;
        off[5] += Amp(0, 2e-5);
        check("agreesWith flags a positive difference", !psi.agreesWith(off));
        off = drawn;
        off[5] += Amp(0, 2e-5);
        check("agreesWith honours tol", psi.agreesWith(off, 1e-4));
        rho.setRandomAmps();
        AmpMatrix m = rho.getAllMatrAmps();
        check("DensityMatrix::agreesWith itself", rho.agreesWith(m));
        m[1][2] -= Amp(0, 1e-3);
        check("DensityMatrix::agreesWith flags a difference", !rho.agreesWith(m));
    }
    comm_end();
    std::printf("drop-in tests: %d trials per op, %d failures\n", TRIALS, g_failures);
    return g_failures == 0 ? 0 : 1;
}
