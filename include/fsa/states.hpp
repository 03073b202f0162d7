// states.hpp -- drop-in for the reference's src/states.hpp.  The amplitudes
// live in HBM behind a libfsa handle; `amps` is a host mirror refreshed by
// getAllVecAmps()/syncFromDevice() and pushed back by syncToDevice() (call it
// after writing `amps` directly).  Geometry fields follow the reference at one
// rank (rank 0 of 1, logNumAmpsPerNode = numQubits, 2N for a density matrix).
#ifndef STATES_HPP
#define STATES_HPP

#include <utility>

#include "bit_maths.hpp"
#include "communication.hpp"
#include "types.hpp"

class StateVector {
public:
    Nat rank;
    Nat numNodes;
    Nat logNumNodes;

    Nat numQubits;
    Index numAmpsPerNode;
    Index logNumAmpsPerNode;

    AmpArray amps;      // host mirror of the device amplitudes
    AmpArray buffer;    // exchange buffer of the multi-rank algorithms; unused at one rank

    fsa_state* handle = nullptr;

    explicit StateVector(Nat numQubits, int precision = FSA_FP64) : StateVector(numQubits, false, precision) {}

    StateVector(const StateVector& o) : StateVector(o.numQubits, o.isDensityMatrix(), o.precision()) {
        fsa_check(fsa_state_copy(handle, o.handle), "StateVector copy");
        amps = o.amps;
    }
    StateVector(StateVector&& o) noexcept { steal(o); }
    StateVector& operator=(StateVector&& o) noexcept {
        if (this != &o) { release(); steal(o); }
        return *this;
    }
    StateVector& operator=(const StateVector& o) {
        if (this != &o) { StateVector t(o); release(); steal(t); }
        return *this;
    }
    ~StateVector() { release(); }

    // test_utilities.hpp:419-435 (one rank: the local amplitudes are all of them)
    AmpArray getAllVecAmps() {
        syncFromDevice();
        return amps;
    }
    // states.hpp:27-30; defined in test_utilities.hpp, as in the reference
    void setRandomAmps();
    void printAmps();
    bool agreesWith(AmpArray allAmps, Real tol = 1E-5);
    void syncFromDevice() {
        amps.resize(numAmpsPerNode);
        fsa_check(fsa_state_download(handle, reinterpret_cast<double*>(amps.data()), numAmpsPerNode), "download");
    }
    void syncToDevice() {
        assert(amps.size() == numAmpsPerNode);
        fsa_check(fsa_state_upload(handle, reinterpret_cast<const double*>(amps.data()), numAmpsPerNode), "upload");
    }
    bool isDensityMatrix() const {
        int dm = 0;
        fsa_state_info(handle, nullptr, &dm, nullptr, nullptr);
        return dm != 0;
    }
    int precision() const {
        int p = 0;
        fsa_state_info(handle, nullptr, nullptr, &p, nullptr);
        return p;
    }

protected:
    StateVector(Nat n, bool dm, int precision) {
        fsa_check(fsa_state_create(n, dm ? 1 : 0, precision, &handle), "StateVector");
        rank = comm_getRank();
        numNodes = comm_getNumNodes();
        logNumNodes = 0;
        numQubits = n;
        logNumAmpsPerNode = dm ? 2 * n : n;
        numAmpsPerNode = powerOf2((Nat)logNumAmpsPerNode);
        amps = AmpArray(numAmpsPerNode, 0);   // the device state starts all-zero too
    }
    explicit StateVector(fsa_state* h) : handle(h) {
        unsigned n = 0;
        int dm = 0;
        uint64_t num = 0;
        fsa_check(fsa_state_info(h, &n, &dm, nullptr, &num), "state info");
        rank = 0;
        numNodes = 1;
        logNumNodes = 0;
        numQubits = n;
        numAmpsPerNode = num;
        logNumAmpsPerNode = dm ? 2 * n : n;
        syncFromDevice();
    }

private:
    void release() {
        if (handle) fsa_state_destroy(handle);
        handle = nullptr;
    }
    void steal(StateVector& o) {
        rank = o.rank; numNodes = o.numNodes; logNumNodes = o.logNumNodes; numQubits = o.numQubits;
        numAmpsPerNode = o.numAmpsPerNode; logNumAmpsPerNode = o.logNumAmpsPerNode;
        amps = std::move(o.amps);
        buffer = std::move(o.buffer);
        handle = o.handle;
        o.handle = nullptr;
    }
};

class DensityMatrix : public StateVector {
public:
    explicit DensityMatrix(Nat numQubits, int precision = FSA_FP64) : StateVector(numQubits, true, precision) {}
    // adopts a handle returned by the library (e.g. a partial trace)
    static DensityMatrix adopt(fsa_state* h) { return DensityMatrix(h); }

    // test_utilities.hpp:490-503: rho[r][c] = amps[r + 2^N c]
    AmpMatrix getAllMatrAmps() {
        AmpArray v = getAllVecAmps();
        Index dim = powerOf2(numQubits);
        AmpMatrix m = getZeroMatrix(dim);
        for (Index r = 0; r < dim; r++)
            for (Index c = 0; c < dim; c++) m[r][c] = v[dim * c + r];
        return m;
    }
    // states.hpp:54-55; defined in test_utilities.hpp
    bool agreesWith(AmpMatrix allAmps, Real tol = 1E-5);

private:
    explicit DensityMatrix(fsa_state* h) : StateVector(h) {}
};

#endif  // STATES_HPP
