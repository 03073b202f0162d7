"""ctypes bindings for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (the C99/OpenMP restatement of the reference's
single-rank loops, see ``fsa_oracle.h`` for the conventions and the D1-D3
contract) and, when it was built, ``oracle/_ref/libfsa_ref.so`` (the unmodified
reference headers behind a C shim).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg import this module.

All state arrays are numpy complex128 (interleaved re/im, same layout as the
reference's ``std::complex<double>``); they are updated in place.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libfsa_ref.so"

_u = C.c_uint
_d = C.c_double
_p = C.c_void_p


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _nats(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(list(v), dtype=np.uint32))


def _mat(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128))


def _check_state(a: np.ndarray):
    if a.dtype != np.complex128 or not a.flags.c_contiguous:
        raise TypeError("oracle states are contiguous complex128 arrays")
    n = int(a.size).bit_length() - 1
    if (1 << n) != a.size:
        raise ValueError("state length must be a power of two")
    return n


class _Lib:
    def __init__(self, path: Path, prefix: str):
        self.lib = C.CDLL(str(path))
        self.prefix = prefix

    def fn(self, name, argtypes, restype=None):
        f = getattr(self.lib, self.prefix + name)
        f.argtypes = argtypes
        f.restype = restype
        return f


class Oracle:
    """Wrapper exposing the oracle's ops with the reference's argument meaning."""

    def __init__(self, path: Path = LIB_PATH, prefix: str = "orc_"):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = _Lib(Path(path), prefix)
        self._L = L
        self.prefix = prefix
        sv_sig = {
            "sv_oneTargGate": [_p, _u, _u, _p],
            "sv_manyCtrlOneTargGate": [_p, _u, _p, _u, _u, _p],
            "sv_swapGate": [_p, _u, _u, _u],
            "sv_manyTargGate": [_p, _u, _p, _u, _p],
            "sv_pauliTensor": [_p, _u, _p, _p, _u],
            "sv_pauliGadget": [_p, _u, _p, _p, _u, _d],
            "sv_phaseGadget": [_p, _u, _p, _u, _d],
            "dm_manyTargGate": [_p, _u, _p, _u, _p],
            "dm_swapGate": [_p, _u, _u, _u],
            "dm_pauliTensor": [_p, _u, _p, _p, _u],
            "dm_pauliGadget": [_p, _u, _p, _p, _u, _d],
            "dm_phaseGadget": [_p, _u, _p, _u, _d],
            "dm_oneQubitDephasing": [_p, _u, _u, _d],
            "dm_twoQubitDephasing": [_p, _u, _u, _u, _d],
            "dm_oneQubitDepolarising": [_p, _u, _u, _d],
            "dm_twoQubitDepolarising": [_p, _u, _u, _u, _d],
            "dm_damping": [_p, _u, _u, _d],
            "dm_expecPauliString": [_p, _u, _p, _u, _p, _p],
            "dm_partialTrace": [_p, _u, _p, _u, _p],
        }
        extra = {
            "dm_krausMap": [_p, _u, _p, _u, _p, _u],
            "dm_twoQubitDepolarising_asShipped": [_p, _u, _u, _u, _d],
            "sv_applyPauliSum": [_p, _p, _u, _p, _u, _p],
            "sv_innerProduct": [_p, _p, _u, _p],
            "sv_pauliMatrixElement": [_p, _p, _u, _p, _p, _u, _p],
            "init_gaussian": [_p, _u, C.c_uint64],
            "num_threads": [],
        }
        self._f = {}
        for name, sig in list(sv_sig.items()) + list(extra.items()):
            try:
                self._f[name] = L.fn(name, sig, C.c_int if name == "num_threads" else None)
            except AttributeError:
                pass

    def has(self, name: str) -> bool:
        return name in self._f

    # ---- state vector ------------------------------------------------------
    def sv_oneTargGate(self, psi, target, gate):
        n = _check_state(psi)
        g = _mat(gate)
        self._f["sv_oneTargGate"](_ptr(psi), n, target, _ptr(g))

    def sv_manyCtrlOneTargGate(self, psi, ctrls, target, gate):
        n = _check_state(psi)
        c, g = _nats(ctrls), _mat(gate)
        self._f["sv_manyCtrlOneTargGate"](_ptr(psi), n, _ptr(c), len(c), target, _ptr(g))

    def sv_swapGate(self, psi, q1, q2):
        n = _check_state(psi)
        self._f["sv_swapGate"](_ptr(psi), n, q1, q2)

    def sv_manyTargGate(self, psi, targets, gate):
        n = _check_state(psi)
        t, g = _nats(targets), _mat(gate)
        self._f["sv_manyTargGate"](_ptr(psi), n, _ptr(t), len(t), _ptr(g))

    def sv_pauliTensor(self, psi, targets, paulis):
        n = _check_state(psi)
        t, p = _nats(targets), _nats(paulis)
        self._f["sv_pauliTensor"](_ptr(psi), n, _ptr(t), _ptr(p), len(t))

    def sv_pauliGadget(self, psi, targets, paulis, theta):
        n = _check_state(psi)
        t, p = _nats(targets), _nats(paulis)
        self._f["sv_pauliGadget"](_ptr(psi), n, _ptr(t), _ptr(p), len(t), float(theta))

    def sv_phaseGadget(self, psi, targets, theta):
        n = _check_state(psi)
        t = _nats(targets)
        self._f["sv_phaseGadget"](_ptr(psi), n, _ptr(t), len(t), float(theta))

    # ---- density matrix (rho given as its 2N-qubit column-major vector) ----
    @staticmethod
    def _nq(rho):
        n2 = _check_state(rho)
        if n2 % 2:
            raise ValueError("density-matrix vectors have an even number of index bits")
        return n2 // 2

    def dm_manyTargGate(self, rho, targets, gate):
        N = self._nq(rho)
        t, g = _nats(targets), _mat(gate)
        self._f["dm_manyTargGate"](_ptr(rho), N, _ptr(t), len(t), _ptr(g))

    def dm_swapGate(self, rho, q1, q2):
        self._f["dm_swapGate"](_ptr(rho), self._nq(rho), q1, q2)

    def dm_pauliTensor(self, rho, targets, paulis):
        N = self._nq(rho)
        t, p = _nats(targets), _nats(paulis)
        self._f["dm_pauliTensor"](_ptr(rho), N, _ptr(t), _ptr(p), len(t))

    def dm_pauliGadget(self, rho, targets, paulis, theta):
        N = self._nq(rho)
        t, p = _nats(targets), _nats(paulis)
        self._f["dm_pauliGadget"](_ptr(rho), N, _ptr(t), _ptr(p), len(t), float(theta))

    def dm_phaseGadget(self, rho, targets, theta):
        N = self._nq(rho)
        t = _nats(targets)
        self._f["dm_phaseGadget"](_ptr(rho), N, _ptr(t), len(t), float(theta))

    def dm_krausMap(self, rho, kraus_ops, targets):
        N = self._nq(rho)
        t = _nats(targets)
        ks = np.ascontiguousarray(np.asarray(kraus_ops, dtype=np.complex128))
        self._f["dm_krausMap"](_ptr(rho), N, _ptr(ks), ks.shape[0], _ptr(t), len(t))

    def dm_oneQubitDephasing(self, rho, q, p):
        self._f["dm_oneQubitDephasing"](_ptr(rho), self._nq(rho), q, float(p))

    def dm_twoQubitDephasing(self, rho, q1, q2, p):
        self._f["dm_twoQubitDephasing"](_ptr(rho), self._nq(rho), q1, q2, float(p))

    def dm_oneQubitDepolarising(self, rho, q, p):
        self._f["dm_oneQubitDepolarising"](_ptr(rho), self._nq(rho), q, float(p))

    def dm_twoQubitDepolarising(self, rho, q1, q2, p):
        self._f["dm_twoQubitDepolarising"](_ptr(rho), self._nq(rho), q1, q2, float(p))

    def dm_twoQubitDepolarising_asShipped(self, rho, q1, q2, p):
        self._f["dm_twoQubitDepolarising_asShipped"](_ptr(rho), self._nq(rho), q1, q2, float(p))

    def dm_damping(self, rho, q, p):
        self._f["dm_damping"](_ptr(rho), self._nq(rho), q, float(p))

    def dm_expecPauliString(self, rho, coeffs, all_paulis) -> complex:
        N = self._nq(rho)
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
        p = _nats(np.asarray(all_paulis).ravel())
        if p.size != c.size * N:
            raise ValueError("allPaulis must hold numTerms * N codes")
        out = np.zeros(2)
        self._f["dm_expecPauliString"](_ptr(rho), N, _ptr(c), c.size, _ptr(p), _ptr(out))
        return complex(out[0], out[1])

    def dm_partialTrace(self, rho, targets) -> np.ndarray:
        N = self._nq(rho)
        t = _nats(targets)
        out = np.zeros(1 << (2 * (N - len(t))), dtype=np.complex128)
        self._f["dm_partialTrace"](_ptr(rho), N, _ptr(t), len(t), _ptr(out))
        return out

    # ---- extensions ------------------------------------------------------
    def init_gaussian(self, n: int, seed: int) -> np.ndarray:
        a = np.empty(1 << n, dtype=np.complex128)
        self._f["init_gaussian"](_ptr(a), n, C.c_uint64(seed))
        return a

    def sv_applyPauliSum(self, psi, coeffs, paulis) -> np.ndarray:
        n = _check_state(psi)
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
        p = _nats(np.asarray(paulis).ravel())
        out = np.empty_like(psi)
        self._f["sv_applyPauliSum"](_ptr(psi), _ptr(out), n, _ptr(c), c.size, _ptr(p))
        return out

    def sv_innerProduct(self, a, b) -> complex:
        n = _check_state(a)
        out = np.zeros(2)
        self._f["sv_innerProduct"](_ptr(a), _ptr(b), n, _ptr(out))
        return complex(out[0], out[1])

    def sv_pauliMatrixElement(self, a, b, targets, paulis) -> complex:
        n = _check_state(a)
        t, p = _nats(targets), _nats(paulis)
        out = np.zeros(2)
        self._f["sv_pauliMatrixElement"](_ptr(a), _ptr(b), n, _ptr(t), _ptr(p), len(t), _ptr(out))
        return complex(out[0], out[1])

    def num_threads(self) -> int:
        return int(self._f["num_threads"]())


_oracle = None
_ref = None


def oracle() -> Oracle:
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle


def reference():
    """The compiled reference (or None if oracle/_ref was not built)."""
    global _ref
    if _ref is None and REF_PATH.exists():
        _ref = Oracle(REF_PATH, prefix="ref_")
        f = _ref._L.lib.ref_set_threads
        f.argtypes = [C.c_int]
        f.restype = C.c_int
        _ref.set_threads = f
    return _ref


def build(quiet: bool = True) -> None:
    """Compile liboracle.so and, if the reference sources are present, oracle/_ref."""
    import subprocess

    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", str(HERE), "-B", "liboracle.so"], check=True, stdout=out)
    if Path(os.environ.get("FSA_REFERENCE_DIR", "/root/reference"), "src").is_dir():
        subprocess.run(["make", "-C", str(HERE), "-B", "ref"], check=True, stdout=out)
        # the reference's own test suite over the drop-in headers (needs libfsa.so)
        subprocess.run(["make", "-C", str(HERE), "suite"], check=True, stdout=out)
