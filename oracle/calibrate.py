"""Calibrate the CPU oracle against the compiled reference (TEST INFRASTRUCTURE).

    python oracle/calibrate.py [--threads T] [--qubits 26] [--dm-qubits 13]

Times each operation whose reference implementation is free of the setBit
defect (so its timing is representative) on the same state size and thread
count, with the oracle (oracle/liboracle.so, operating in place) and with the
unmodified reference (oracle/_ref/libfsa_ref.so, timed inside the library on a
persistent state so no copies are counted).  Prints one markdown table row per
operation: seconds per call for each, and the ratio.  Needs oracle/_ref, which
is built only where the reference sources are present.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import oracle  # noqa: E402


def _ref_timer():
    ref = oracle.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libfsa_ref.so is not built (make -C oracle ref)")
    L = ref._L.lib
    L.ref_time_state_new.argtypes = [C.c_uint, C.c_int]
    L.ref_time_state_new.restype = C.c_void_p
    L.ref_time_state_free.argtypes = [C.c_void_p, C.c_int]
    L.ref_time_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int]
    L.ref_time_op.restype = C.c_double
    return ref, L


def best_of(fn, reps):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--qubits", type=int, default=26)
    ap.add_argument("--dm-qubits", type=int, default=13)
    ap.add_argument("--expec-qubits", type=int, default=11)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    orc = oracle.oracle()
    ref, L = _ref_timer()
    ref.set_threads(a.threads)
    rng = np.random.default_rng(1)
    g = np.linalg.svd(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    gd = np.ascontiguousarray(g, dtype=np.complex128)
    print(f"threads {a.threads} (oracle reports {orc.num_threads()})")
    print("| op | size | oracle s/call | reference s/call | reference / oracle |")
    print("|---|---|---|---|---|")

    def row(name, size, t_orc, t_ref):
        print(f"| {name} | {size} | {t_orc:.4f} | {t_ref:.4f} | {t_ref / t_orc:.2f} |", flush=True)

    n = a.qubits
    psi = orc.init_gaussian(n, 3)
    sv = L.ref_time_state_new(n, 0)
    cases = [
        ("oneTargGate (t=3)", 0, [3], lambda: orc.sv_oneTargGate(psi, 3, g)),
        ("oneTargGate (t=20)", 0, [20], lambda: orc.sv_oneTargGate(psi, 20, g)),
        ("manyCtrlOneTargGate (2 controls)", 1, [7, 2, 15], lambda: orc.sv_manyCtrlOneTargGate(psi, [2, 15], 7, g)),
        ("swapGate", 2, [4, 19], lambda: orc.sv_swapGate(psi, 4, 19)),
        ("phaseGadget (4 qubits)", 3, [1, 8, 13, 22], lambda: orc.sv_phaseGadget(psi, [1, 8, 13, 22], 0.3)),
    ]
    for name, code, q, f in cases:
        qa = np.array(q + [0] * (4 - len(q)), np.uint32)
        t_ref = min(L.ref_time_op(sv, 0, code, qa.ctypes.data_as(C.c_void_p), gd.ctypes.data_as(C.c_void_p), 0.3, 1)
                    for _ in range(a.reps))
        row(name, f"{n} qubits", best_of(f, a.reps), t_ref)
    L.ref_time_state_free(sv, 0)
    del psi

    N = a.dm_qubits
    rho = orc.init_gaussian(2 * N, 4)
    dm = L.ref_time_state_new(N, 1)
    cases = [
        ("oneQubitDephasing", 10, [3], lambda: orc.dm_oneQubitDephasing(rho, 3, 0.05)),
        ("twoQubitDephasing", 11, [2, 9], lambda: orc.dm_twoQubitDephasing(rho, 2, 9, 0.05)),
        ("oneQubitDepolarising", 12, [5], lambda: orc.dm_oneQubitDepolarising(rho, 5, 0.05)),
        ("damping", 13, [6], lambda: orc.dm_damping(rho, 6, 0.05)),
    ]
    for name, code, q, f in cases:
        qa = np.array(q + [0] * (4 - len(q)), np.uint32)
        t_ref = min(L.ref_time_op(dm, 1, code, qa.ctypes.data_as(C.c_void_p), None, 0.05, 1) for _ in range(a.reps))
        row(name, f"N = {N} (2^{2 * N} amplitudes)", best_of(f, a.reps), t_ref)
    L.ref_time_state_free(dm, 1)
    del rho

    N = a.expec_qubits
    rho = orc.init_gaussian(2 * N, 5)
    dm = L.ref_time_state_new(N, 1)
    qa = np.array([0, 3, 5, 7], np.uint32)
    codes = np.zeros(N, np.uint32)
    for i, q in enumerate(qa):
        codes[q] = 1 + i % 3
    t_ref = L.ref_time_op(dm, 1, 14, qa.ctypes.data_as(C.c_void_p), None, 0.7, 1)
    t_orc = best_of(lambda: orc.dm_expecPauliString(rho, [0.7], codes), a.reps)
    row("expecPauliString (1 term)", f"N = {N}", t_orc, t_ref)
    L.ref_time_state_free(dm, 1)


if __name__ == "__main__":
    main()
