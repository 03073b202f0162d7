"""fsa_amd -- Python face of the MI355X full-state gate engine (libfsa.so).

Mirrors the reference's C++ API for one rank: ``StateVector``/``DensityMatrix``
(reference src/states.hpp:13-69) and the ``distributed_statevector_*`` /
``distributed_densitymatrix_*`` functions (src/distributed_statevector.hpp,
src/distributed_densitymatrix.hpp), with the same argument meaning.  Every call
goes through the C ABI declared in include/fsa.h; there is no CPU fallback:
if libfsa.so is missing or no GPU is present the calls raise.

Errors: the reference asserts; here a non-zero status raises ``FsaError``
carrying the library's code and message.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Iterable, Sequence

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libfsa.so"

FSA_FP64, FSA_FP32 = 0, 1
FSA_EXEC_EAGER, FSA_EXEC_FUSED = 0, 1
I, X, Y, Z = 0, 1, 2, 3          # reference src/types.hpp:16-18

ERR_NAMES = {1: "FSA_ERR_ARG", 2: "FSA_ERR_QUBIT", 3: "FSA_ERR_DUPLICATE", 4: "FSA_ERR_SIZE",
             5: "FSA_ERR_OOM", 6: "FSA_ERR_HIP", 7: "FSA_ERR_KIND", 8: "FSA_ERR_INTERNAL"}


class FsaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_u, _i, _d, _p, _u64 = C.c_uint, C.c_int, C.c_double, C.c_void_p, C.c_uint64
_P = C.POINTER


def lib() -> C.CDLL:
    """The loaded libfsa.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        L.fsa_last_error.restype = C.c_char_p
        L.fsa_comm_getRank.restype = _u
        L.fsa_comm_getNumNodes.restype = _u
        _lib = L
    return _lib


def _check(code: int):
    if code != 0:
        raise FsaError(code, lib().fsa_last_error().decode())


def _nats(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(list(v) if not isinstance(v, np.ndarray) else v, dtype=np.uint32))


def _mat(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _precision(p) -> int:
    if p in (FSA_FP64, "fp64", "f64", np.complex128, "double"):
        return FSA_FP64
    if p in (FSA_FP32, "fp32", "f32", np.complex64, "float"):
        return FSA_FP32
    raise ValueError(f"unknown precision {p!r}")


# ---- communication (reference src/communication.hpp; one rank) -------------

def comm_init():
    _check(lib().fsa_comm_init())


def comm_end():
    _check(lib().fsa_comm_end())


def comm_getRank() -> int:
    return int(lib().fsa_comm_getRank())


def comm_getNumNodes() -> int:
    return int(lib().fsa_comm_getNumNodes())


def comm_synch():
    _check(lib().fsa_comm_synch())


def set_two_qubit_depolarising_as_shipped(enable: bool):
    _check(lib().fsa_set_two_qubit_depolarising_as_shipped(1 if enable else 0))


# ---- states (reference src/states.hpp) ---------------------------------------

class StateVector:
    """A statevector of ``numQubits`` qubits held in HBM (reference states.hpp:13-49).

    Geometry fields follow the reference at one rank: rank 0, numNodes 1,
    logNumAmpsPerNode = numQubits.  Amplitudes live on the GPU; ``amps`` reads
    them back (complex128) and assigning it uploads.
    """

    _is_dm = False

    def __init__(self, numQubits: int, precision="fp64", mode: int = FSA_EXEC_FUSED, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            _check(lib().fsa_state_create(_u(numQubits), _i(1 if self._is_dm else 0), _i(_precision(precision)),
                                          C.byref(self._h)))
        nq, dm, prec, num = _u(), _i(), _i(), _u64()
        _check(lib().fsa_state_info(self._h, C.byref(nq), C.byref(dm), C.byref(prec), C.byref(num)))
        self.numQubits = int(nq.value)
        self.precision = int(prec.value)
        self.numAmpsPerNode = int(num.value)
        self.logNumAmpsPerNode = self.numAmpsPerNode.bit_length() - 1
        self.rank, self.numNodes, self.logNumNodes = 0, 1, 0
        if mode != FSA_EXEC_FUSED:
            self.set_exec_mode(mode)

    # lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fsa_state_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_exec_mode(self, mode: int):
        _check(lib().fsa_state_set_exec_mode(self._h, _i(mode)))

    def flush(self):
        _check(lib().fsa_state_flush(self._h))

    def sync(self):
        _check(lib().fsa_state_sync(self._h))

    # amplitudes
    def upload(self, amps):
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128).ravel())
        _check(lib().fsa_state_upload(self._h, _ptr(a), _u64(a.size)))

    def download(self) -> np.ndarray:
        out = np.empty(self.numAmpsPerNode, dtype=np.complex128)
        _check(lib().fsa_state_download(self._h, _ptr(out), _u64(out.size)))
        return out

    @property
    def amps(self) -> np.ndarray:
        return self.download()

    @amps.setter
    def amps(self, value):
        self.upload(value)

    def getAllVecAmps(self) -> np.ndarray:          # test_utilities.hpp:419-435
        return self.download()

    def init_zero(self):
        _check(lib().fsa_state_init_zero(self._h))

    def init_gaussian(self, seed: int):
        _check(lib().fsa_state_init_gaussian(self._h, _u64(seed)))

    def renormalise(self):
        """Scale to unit norm (FP64 sum on the device); statevectors only."""
        _check(lib().fsa_statevector_renormalise(self._h))

    def norm2(self) -> float:
        out = _d()
        _check(lib().fsa_state_norm2(self._h, C.byref(out)))
        return out.value

    def copy_from(self, other: "StateVector"):
        _check(lib().fsa_state_copy(self._h, other._h))

    def device_ptr(self) -> int:
        p = C.c_void_p()
        _check(lib().fsa_state_device_ptr(self._h, C.byref(p)))
        return int(p.value or 0)

    def stream(self) -> int:
        p = C.c_void_p()
        _check(lib().fsa_state_stream(self._h, C.byref(p)))
        return int(p.value or 0)


class DensityMatrix(StateVector):
    """A density matrix of ``numQubits`` qubits: the column-major 2N-qubit vector
    amps[r + 2^N c] = rho[r][c] (reference states.hpp:53-69, test_utilities.hpp:500)."""

    _is_dm = True

    def getAllMatrAmps(self) -> np.ndarray:         # test_utilities.hpp:490-503
        v = self.download()
        D = 1 << self.numQubits
        return v.reshape(D, D).T.copy()


# ---- statevector API (reference src/distributed_statevector.hpp) -------------

def distributed_statevector_oneTargGate(psi: StateVector, target: int, gate):
    g = _mat(gate)
    _check(lib().fsa_statevector_oneTargGate(psi._h, _u(target), _ptr(g)))


def distributed_statevector_manyCtrlOneTargGate(psi: StateVector, controls: Sequence[int], target: int, gate):
    c, g = _nats(controls), _mat(gate)
    _check(lib().fsa_statevector_manyCtrlOneTargGate(psi._h, _ptr(c), _u(c.size), _u(target), _ptr(g)))


def distributed_statevector_swapGate(psi: StateVector, qb1: int, qb2: int):
    _check(lib().fsa_statevector_swapGate(psi._h, _u(qb1), _u(qb2)))


def distributed_statevector_manyTargGate(psi: StateVector, targets: Sequence[int], gate):
    t, g = _nats(targets), _mat(gate)
    _check(lib().fsa_statevector_manyTargGate(psi._h, _ptr(t), _u(t.size), _ptr(g)))


def distributed_statevector_pauliTensor(psi: StateVector, targets, paulis):
    t, p = _nats(targets), _nats(paulis)
    _check(lib().fsa_statevector_pauliTensor(psi._h, _ptr(t), _ptr(p), _u(t.size)))


def distributed_statevector_pauliGadget(psi: StateVector, targets, paulis, theta: float):
    t, p = _nats(targets), _nats(paulis)
    _check(lib().fsa_statevector_pauliGadget(psi._h, _ptr(t), _ptr(p), _u(t.size), _d(theta)))


def distributed_statevector_phaseGadget(psi: StateVector, targets, theta: float):
    t = _nats(targets)
    _check(lib().fsa_statevector_phaseGadget(psi._h, _ptr(t), _u(t.size), _d(theta)))


# ---- density-matrix API (reference src/distributed_densitymatrix.hpp) --------

def distributed_densitymatrix_manyTargGate(rho: DensityMatrix, targets, gate):
    t, g = _nats(targets), _mat(gate)
    _check(lib().fsa_densitymatrix_manyTargGate(rho._h, _ptr(t), _u(t.size), _ptr(g)))


def distributed_densitymatrix_swapGate(rho: DensityMatrix, qb1: int, qb2: int):
    _check(lib().fsa_densitymatrix_swapGate(rho._h, _u(qb1), _u(qb2)))


def distributed_densitymatrix_pauliTensor(rho: DensityMatrix, targets, paulis):
    t, p = _nats(targets), _nats(paulis)
    _check(lib().fsa_densitymatrix_pauliTensor(rho._h, _ptr(t), _ptr(p), _u(t.size)))


def distributed_densitymatrix_pauliGadget(rho: DensityMatrix, targets, paulis, theta: float):
    t, p = _nats(targets), _nats(paulis)
    _check(lib().fsa_densitymatrix_pauliGadget(rho._h, _ptr(t), _ptr(p), _u(t.size), _d(theta)))


def distributed_densitymatrix_phaseGadget(rho: DensityMatrix, targets, theta: float):
    t = _nats(targets)
    _check(lib().fsa_densitymatrix_phaseGadget(rho._h, _ptr(t), _u(t.size), _d(theta)))


def distributed_densitymatrix_krausMap(rho: DensityMatrix, krausOps, targets):
    t = _nats(targets)
    ks = np.ascontiguousarray(np.asarray(krausOps, dtype=np.complex128))
    _check(lib().fsa_densitymatrix_krausMap(rho._h, _ptr(ks), _u(ks.shape[0]), _ptr(t), _u(t.size)))


def distributed_densitymatrix_oneQubitDephasing(rho: DensityMatrix, qb: int, prob: float):
    _check(lib().fsa_densitymatrix_oneQubitDephasing(rho._h, _u(qb), _d(prob)))


def distributed_densitymatrix_twoQubitDephasing(rho: DensityMatrix, qb1: int, qb2: int, prob: float):
    _check(lib().fsa_densitymatrix_twoQubitDephasing(rho._h, _u(qb1), _u(qb2), _d(prob)))


def distributed_densitymatrix_oneQubitDepolarising(rho: DensityMatrix, qb: int, prob: float):
    _check(lib().fsa_densitymatrix_oneQubitDepolarising(rho._h, _u(qb), _d(prob)))


def distributed_densitymatrix_twoQubitDepolarising(rho: DensityMatrix, qb1: int, qb2: int, prob: float):
    _check(lib().fsa_densitymatrix_twoQubitDepolarising(rho._h, _u(qb1), _u(qb2), _d(prob)))


def distributed_densitymatrix_damping(rho: DensityMatrix, qb: int, prob: float):
    _check(lib().fsa_densitymatrix_damping(rho._h, _u(qb), _d(prob)))


def distributed_densitymatrix_expecPauliString(rho: DensityMatrix, coeffs, allPaulis) -> complex:
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    p = _nats(np.asarray(allPaulis).ravel())
    out = np.zeros(2)
    _check(lib().fsa_densitymatrix_expecPauliString(rho._h, _ptr(c), _u(c.size), _ptr(p), _ptr(out)))
    return complex(out[0], out[1])


def distributed_densitymatrix_partialTrace(rho: DensityMatrix, targets) -> DensityMatrix:
    t = _nats(targets)
    h = C.c_void_p()
    _check(lib().fsa_densitymatrix_partialTrace(rho._h, _ptr(t), _u(t.size), C.byref(h)))
    return DensityMatrix(0, _handle=h)


# ---- extensions ---------------------------------------------------------------

def statevector_expecPauliSum(psi: StateVector, coeffs, paulis) -> complex:
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    p = _nats(np.asarray(paulis).ravel())
    out = np.zeros(2)
    _check(lib().fsa_statevector_expecPauliSum(psi._h, _ptr(c), _u(c.size), _ptr(p), _ptr(out)))
    return complex(out[0], out[1])


def statevector_applyPauliSum(psi: StateVector, coeffs, paulis, out: StateVector):
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    p = _nats(np.asarray(paulis).ravel())
    _check(lib().fsa_statevector_applyPauliSum(psi._h, _ptr(c), _u(c.size), _ptr(p), out._h))


def statevector_innerProduct(a: StateVector, b: StateVector) -> complex:
    out = np.zeros(2)
    _check(lib().fsa_statevector_innerProduct(a._h, b._h, _ptr(out)))
    return complex(out[0], out[1])


# ---- circuits, adjoint gradient, profiling ------------------------------------

class Circuit:
    """A recorded statevector circuit with a cached fused schedule (C ABI fsa_circuit_*).

    ``ops`` is a list of (name, args) pairs as produced by ``circuits.py``;
    pauliGadget / phaseGadget angles become parameters in order.
    """

    def __init__(self, ops: Iterable = ()):
        self._h = C.c_void_p()
        _check(lib().fsa_circuit_create(C.byref(self._h)))
        for name, args in ops:
            self.add(name, *args)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fsa_circuit_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, name: str, *args):
        L = lib()
        if name == "oneTargGate":
            t, g = args
            gm = _mat(g)
            _check(L.fsa_circuit_oneTargGate(self._h, _u(t), _ptr(gm)))
        elif name == "manyCtrlOneTargGate":
            c, t, g = args
            cc, gm = _nats(c), _mat(g)
            _check(L.fsa_circuit_manyCtrlOneTargGate(self._h, _ptr(cc), _u(cc.size), _u(t), _ptr(gm)))
        elif name == "swapGate":
            _check(L.fsa_circuit_swapGate(self._h, _u(args[0]), _u(args[1])))
        elif name == "manyTargGate":
            t, gm = _nats(args[0]), _mat(args[1])
            _check(L.fsa_circuit_manyTargGate(self._h, _ptr(t), _u(t.size), _ptr(gm)))
        elif name == "pauliTensor":
            t, p = _nats(args[0]), _nats(args[1])
            _check(L.fsa_circuit_pauliTensor(self._h, _ptr(t), _ptr(p), _u(t.size)))
        elif name == "pauliGadget":
            t, p = _nats(args[0]), _nats(args[1])
            _check(L.fsa_circuit_pauliGadget(self._h, _ptr(t), _ptr(p), _u(t.size), _d(args[2])))
        elif name == "phaseGadget":
            t = _nats(args[0])
            _check(L.fsa_circuit_phaseGadget(self._h, _ptr(t), _u(t.size), _d(args[1])))
        else:
            raise ValueError(f"unknown circuit op {name}")

    @property
    def num_ops(self) -> int:
        v = _u()
        _check(lib().fsa_circuit_num_ops(self._h, C.byref(v)))
        return int(v.value)

    @property
    def num_params(self) -> int:
        v = _u()
        _check(lib().fsa_circuit_num_params(self._h, C.byref(v)))
        return int(v.value)

    def set_params(self, thetas):
        t = np.ascontiguousarray(np.asarray(thetas, dtype=np.float64))
        _check(lib().fsa_circuit_set_params(self._h, _ptr(t), _u(t.size)))

    def apply(self, psi: StateVector):
        _check(lib().fsa_circuit_apply(self._h, psi._h))

    def num_passes(self, numQubits: int, precision="fp64") -> int:
        v = _u()
        _check(lib().fsa_circuit_num_passes(self._h, _u(numQubits), _i(_precision(precision)), C.byref(v)))
        return int(v.value)


def adjoint_gradient(circuit: Circuit, psi0: StateVector, coeffs, paulis):
    """<H> and d<H>/d theta for every gadget parameter (adjoint method)."""
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    p = _nats(np.asarray(paulis).ravel())
    expec = _d()
    grads = np.zeros(circuit.num_params)
    _check(lib().fsa_adjoint_gradient(circuit._h, psi0._h, _ptr(c), _u(c.size), _ptr(p), C.byref(expec),
                                      _ptr(grads)))
    return expec.value, grads


class Batch:
    """Many independent circuits, each on its own statevector from |0..0> with
    its own observable; ``run()`` returns every <H> and every gadget gradient
    from one persistent-kernel launch (C ABI fsa_batch_*)."""

    def __init__(self, precision="fp64"):
        self._h = C.c_void_p()
        _check(lib().fsa_batch_create(_i(_precision(precision)), C.byref(self._h)))
        self.sizes = []            # parameter count per circuit

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fsa_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, circuit: Circuit, numQubits: int, coeffs, paulis) -> int:
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
        p = _nats(np.asarray(paulis).ravel())
        if p.size != c.size * numQubits:
            raise ValueError("paulis must hold numTerms * numQubits codes")
        idx = _u()
        _check(lib().fsa_batch_add(self._h, circuit._h, _u(numQubits), _ptr(c), _u(c.size), _ptr(p), C.byref(idx)))
        self.sizes.append(circuit.num_params)
        return int(idx.value)

    @property
    def num_circuits(self) -> int:
        n, p = _u(), _u()
        _check(lib().fsa_batch_info(self._h, C.byref(n), C.byref(p)))
        return int(n.value)

    @property
    def num_params(self) -> int:
        n, p = _u(), _u()
        _check(lib().fsa_batch_info(self._h, C.byref(n), C.byref(p)))
        return int(p.value)

    def set_params(self, index: int, thetas):
        t = np.ascontiguousarray(np.asarray(thetas, dtype=np.float64))
        _check(lib().fsa_batch_set_params(self._h, _u(index), _ptr(t), _u(t.size)))

    def set_config(self, workgroup_size: int, tile_bits: int, low_bits: int):
        _check(lib().fsa_batch_set_config(self._h, _i(workgroup_size), _i(tile_bits), _i(low_bits)))

    @property
    def config(self):
        a, b, c = _i(), _i(), _i()
        _check(lib().fsa_batch_get_config(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def autotune(self, sample: int = 0) -> float:
        ms = _d()
        _check(lib().fsa_batch_autotune(self._h, _u(sample), C.byref(ms)))
        return float(ms.value)

    def run(self):
        """(expecs[num_circuits], [grads of circuit 0, grads of circuit 1, ...])"""
        e = np.zeros(self.num_circuits)
        g = np.zeros(max(1, self.num_params))
        _check(lib().fsa_batch_run(self._h, _ptr(e), _ptr(g)))
        out, k = [], 0
        for n in self.sizes:
            out.append(g[k:k + n].copy())
            k += n
        return e, out


class _ProfEntry(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("count", C.c_uint64), ("total_ms", C.c_double)]


def profile_enable(enable: bool = True):
    _check(lib().fsa_profile_enable(_i(1 if enable else 0)))


def profile_read() -> dict:
    """{kernel kind: (launch count, total milliseconds)} since profile_enable."""
    num = _u()
    _check(lib().fsa_profile_read(None, _u(0), C.byref(num)))
    arr = (_ProfEntry * max(1, num.value))()
    _check(lib().fsa_profile_read(arr, _u(num.value), C.byref(num)))
    return {arr[i].name.decode(): (int(arr[i].count), float(arr[i].total_ms)) for i in range(num.value)}


def apply_ops(state: StateVector, ops, dm: bool = False):
    """Apply a (name, args) list through the reference-named API functions."""
    prefix = "distributed_densitymatrix_" if dm else "distributed_statevector_"
    g = globals()
    for name, args in ops:
        g[prefix + name](state, *args)
