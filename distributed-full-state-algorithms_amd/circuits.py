"""Seeded synthetic circuits for the benchmark configurations.

A circuit is a list of ``(name, args)`` pairs where ``name`` is the reference's
operation name without its ``distributed_statevector_`` / ``distributed_densitymatrix_``
prefix and ``args`` are that function's arguments after the state
(reference src/distributed_statevector.hpp, src/distributed_densitymatrix.hpp).
Both the GPU engine and the CPU oracle consume the same list.

Random laws (numpy PCG64, seed given):
  * Haar U(d): QR of a complex Ginibre matrix with diag(R) phases removed;
  * angles theta ~ U(0, 2 pi); probabilities ~ U(0, 0.1);
  * qubits uniform over the register, distinct within one op.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2 * np.pi


def haar(d: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def _qubits(rng, n, k):
    return [int(x) for x in rng.choice(n, k, replace=False)]


def sv_mix(n: int, num_gates: int, seed: int, extended: bool) -> list:
    """Statevector gate mix.

    Base mix (12-qubit, 200-gate configuration): oneTargGate (Haar 2x2, uniform
    target); manyCtrlOneTargGate (2 controls, Haar 2x2); swapGate; pauliGadget
    on 3 qubits with Paulis in {X, Y, Z} and theta ~ U(0, 2pi).
    Extended mix (30-qubit, 1000-gate configuration) adds manyTargGate with
    k in {2,3,4,5} (Haar 2^k), phaseGadget on 4 qubits and pauliTensor on 6 qubits.
    """
    rng = np.random.default_rng(seed)
    kinds = ["one", "ctrl", "swap", "gadget"] + (["many", "phase", "tensor"] if extended else [])
    ops = []
    for _ in range(num_gates):
        kind = kinds[int(rng.integers(len(kinds)))]
        if kind == "one":
            ops.append(("oneTargGate", (int(rng.integers(n)), haar(2, rng))))
        elif kind == "ctrl":
            q = _qubits(rng, n, 3)
            ops.append(("manyCtrlOneTargGate", (q[1:], q[0], haar(2, rng))))
        elif kind == "swap":
            q = _qubits(rng, n, 2)
            ops.append(("swapGate", (q[0], q[1])))
        elif kind == "gadget":
            q = _qubits(rng, n, 3)
            p = [int(x) for x in rng.integers(1, 4, 3)]
            ops.append(("pauliGadget", (q, p, float(rng.uniform(0, TWO_PI)))))
        elif kind == "many":
            k = int(rng.integers(2, 6))
            ops.append(("manyTargGate", (_qubits(rng, n, k), haar(1 << k, rng))))
        elif kind == "phase":
            ops.append(("phaseGadget", (_qubits(rng, n, 4), float(rng.uniform(0, TWO_PI)))))
        else:
            q = _qubits(rng, n, 6)
            p = [int(x) for x in rng.integers(1, 4, 6)]
            ops.append(("pauliTensor", (q, p)))
    return ops


def config1(seed: int = 20261203) -> list:
    """12-qubit statevector, 200 gates, base mix."""
    return sv_mix(12, 200, seed, extended=False)


def config2(seed: int = 20261204, n: int = 30, num_gates: int = 1000) -> list:
    """30-qubit statevector, 1000 gates, extended mix."""
    return sv_mix(n, num_gates, seed, extended=True)


def dm_mix(N: int, num_ops: int, seed: int) -> list:
    """Density-matrix sequence: manyTargGate k in {1,2,3} (Haar); one/two-qubit
    dephasing and depolarising, damping; prob ~ U(0, 0.1)."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(num_ops):
        kind = int(rng.integers(6))
        p = float(rng.uniform(0, 0.1))
        if kind == 0:
            k = int(rng.integers(1, 4))
            ops.append(("manyTargGate", (_qubits(rng, N, k), haar(1 << k, rng))))
        elif kind == 1:
            ops.append(("oneQubitDephasing", (int(rng.integers(N)), p)))
        elif kind == 2:
            q = _qubits(rng, N, 2)
            ops.append(("twoQubitDephasing", (q[0], q[1], p)))
        elif kind == 3:
            ops.append(("oneQubitDepolarising", (int(rng.integers(N)), p)))
        elif kind == 4:
            q = _qubits(rng, N, 2)
            ops.append(("twoQubitDepolarising", (q[0], q[1], p)))
        else:
            ops.append(("damping", (int(rng.integers(N)), p)))
    return ops


def config3(seed: int = 20261205, N: int = 14, num_ops: int = 400):
    """14-qubit density matrix: 400 ops, then 64 Pauli strings (coefficients
    N(0,1)) for expecPauliString and 4 qubits for partialTrace."""
    ops = dm_mix(N, num_ops, seed)
    rng = np.random.default_rng(seed + 1)
    paulis = rng.integers(0, 4, 64 * N)
    coeffs = rng.normal(size=64)
    trace_targets = _qubits(rng, N, 4)
    return ops, coeffs, paulis, trace_targets


def random_mixed_state(N: int, seed: int, rank: int = 8) -> np.ndarray:
    """rho = sum_i w_i |g_i><g_i|, w ~ Dirichlet(1), g_i normalised Gaussians;
    returned as the column-major 2N-qubit vector."""
    rng = np.random.default_rng(seed)
    D = 1 << N
    w = rng.dirichlet(np.ones(rank))
    G = np.empty((rank, D), dtype=complex)
    for i in range(rank):
        g = rng.normal(size=D) + 1j * rng.normal(size=D)
        G[i] = g / np.linalg.norm(g)
    # column-major vector: amps[r + D c] = rho[r, c] = sum_i w_i g_i[r] conj(g_i[c])
    return ((G.conj().T * w) @ G).reshape(-1)


def config4(n: int = 24, layers: int = 20, seed: int = 20261206, num_terms: int = 100):
    """Layered ZZ + X gadget circuit from |0..0>: per layer a ZZ gadget on each of
    the n-1 nearest-neighbour pairs, then an X gadget on every qubit; angles
    ~ U(0, 2pi).  Observable: num_terms Pauli strings of weight 1..4 on random
    qubits, codes in {X, Y, Z}, coefficients N(0,1) (term-major n codes each)."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        for q in range(n - 1):
            ops.append(("pauliGadget", ([q, q + 1], [3, 3], float(rng.uniform(0, TWO_PI)))))
        for q in range(n):
            ops.append(("pauliGadget", ([q], [1], float(rng.uniform(0, TWO_PI)))))
    paulis = np.zeros((num_terms, n), dtype=np.uint32)
    for t in range(num_terms):
        w = int(rng.integers(1, 5))
        qs = _qubits(rng, n, w)
        paulis[t, qs] = rng.integers(1, 4, w)
    coeffs = rng.normal(size=num_terms)
    return ops, coeffs, paulis


def haar_stack(count: int, d: int, rng: np.random.Generator) -> np.ndarray:
    """count independent Haar U(d) matrices (the law of haar(), vectorised)."""
    z = (rng.normal(size=(count, d, d)) + 1j * rng.normal(size=(count, d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    dg = np.diagonal(r, axis1=1, axis2=2)
    return q * (dg / np.abs(dg))[:, None, :]


def config5(seed: int = 20261207, num: int = 4096, nmin: int = 10, nmax: int = 20, num_terms: int = 20):
    """Batch of independent circuits: n = round(exp(U(ln nmin, ln nmax))),
    50-300 gates each, uniform over oneTargGate (Haar 2x2), manyCtrlOneTargGate
    (1-3 controls, Haar 2x2), manyTargGate (k = 2 or 3, Haar) and pauliGadget
    (1-4 qubits, Paulis in {X, Y, Z}, theta ~ U(0, 2pi)); each circuit has its
    own num_terms-string observable (weight 1-4, coefficients N(0,1)).
    Returns [(n, ops, coeffs, paulis)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(num):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        g = int(rng.integers(50, 301))
        kinds = rng.integers(0, 4, g)
        ks = rng.integers(2, 4, g)                     # manyTargGate sizes
        u2 = haar_stack(g, 2, rng)
        u4 = haar_stack(g, 4, rng)
        u8 = haar_stack(g, 8, rng)
        perms = np.argsort(rng.random((g, n)), axis=1)  # distinct qubits per op
        ctrl = rng.integers(1, 4, g)
        wid = rng.integers(1, 5, g)
        codes = rng.integers(1, 4, (g, 4))
        angles = rng.uniform(0, TWO_PI, g)
        ops = []
        for i in range(g):
            q = [int(x) for x in perms[i]]
            k = int(kinds[i])
            if k == 0:
                ops.append(("oneTargGate", (q[0], u2[i])))
            elif k == 1:
                c = min(int(ctrl[i]), n - 1)
                ops.append(("manyCtrlOneTargGate", (q[1:1 + c], q[0], u2[i])))
            elif k == 2:
                ops.append(("manyTargGate", (q[:int(ks[i])], u4[i] if ks[i] == 2 else u8[i])))
            else:
                w = int(wid[i])
                ops.append(("pauliGadget", (q[:w], [int(x) for x in codes[i, :w]], float(angles[i]))))
        paulis = np.zeros((num_terms, n), dtype=np.uint32)
        for t in range(num_terms):
            w = int(rng.integers(1, 5))
            paulis[t, _qubits(rng, n, w)] = rng.integers(1, 4, w)
        out.append((n, ops, rng.normal(size=num_terms), paulis))
    return out


def algorithmic_bytes(name: str, args, nq: int, amp_bytes: int = 16, dm: bool = False) -> int:
    """Bytes an op must move: 2 x amp size x amplitudes it can change.

    nq is the number of index bits (n for a statevector, 2N for a density matrix).
    """
    full = 2 * amp_bytes * (1 << nq)
    if dm:
        if name == "oneQubitDephasing":
            return full // 2
        if name == "twoQubitDephasing":
            return full * 3 // 4
        return full
    if name == "manyCtrlOneTargGate":
        return full >> len(args[0])
    if name == "swapGate":
        return full // 2
    return full


# ---- on-disk circuit format ---------------------------------------------------
# One .npz file (numpy's zip of .npy arrays; read with allow_pickle=False):
#   "format"      -> "fsa-circuit-1"
#   "header"      -> JSON text: {"num_qubits": n, "density_matrix": bool,
#                    "ops": [{"name": ..., "qubits": [...], "controls": [...],
#                             "paulis": [...], "theta": x, "prob": p,
#                             "matrix": index, "kraus": index}, ...],
#                    "observable": {"coeffs": index, "paulis": index} | null}
#   "m<i>"        -> complex128 matrices (gate matrices, Kraus stacks), by index
#   "coeffs<i>", "paulis<i>" -> observable arrays
# Names and argument meaning are the reference's operation names
# (src/distributed_statevector.hpp, src/distributed_densitymatrix.hpp).

_ARGS = {
    "oneTargGate": ("qubits", "matrix"), "manyCtrlOneTargGate": ("controls", "qubits", "matrix"),
    "swapGate": ("qubits",), "manyTargGate": ("qubits", "matrix"), "pauliTensor": ("qubits", "paulis"),
    "pauliGadget": ("qubits", "paulis", "theta"), "phaseGadget": ("qubits", "theta"),
    "krausMap": ("kraus", "qubits"), "oneQubitDephasing": ("qubits", "prob"),
    "twoQubitDephasing": ("qubits", "prob"), "oneQubitDepolarising": ("qubits", "prob"),
    "twoQubitDepolarising": ("qubits", "prob"), "damping": ("qubits", "prob"),
}


def save_circuit(path, num_qubits: int, ops, density_matrix: bool = False, coeffs=None, paulis=None) -> None:
    """Write a (name, args) op list, and optionally an observable, to `path` (.npz)."""
    import json

    arrays, recs = {}, []
    for name, args in ops:
        if name not in _ARGS:
            raise ValueError(f"unknown operation {name}")
        rec = {"name": name}
        a = list(args)
        fields = _ARGS[name]
        if name in ("swapGate",):
            rec["qubits"] = [int(a[0]), int(a[1])]
        elif name in ("oneTargGate",):
            rec["qubits"] = [int(a[0])]
        elif name in ("twoQubitDephasing", "twoQubitDepolarising"):
            rec["qubits"] = [int(a[0]), int(a[1])]
            rec["prob"] = float(a[2])
        elif name in ("oneQubitDephasing", "oneQubitDepolarising", "damping"):
            rec["qubits"] = [int(a[0])]
            rec["prob"] = float(a[1])
        else:
            for f, v in zip(fields, a):
                if f in ("qubits", "controls", "paulis"):
                    rec[f] = [int(x) for x in (v if not np.isscalar(v) else [v])]
                elif f == "theta":
                    rec[f] = float(v)
        for f, v in zip(fields, a):
            if f in ("matrix", "kraus"):
                key = f"m{len(arrays)}"
                arrays[key] = np.ascontiguousarray(np.asarray(v, dtype=np.complex128))
                rec[f] = key
        recs.append(rec)
    obs = None
    if coeffs is not None:
        arrays["coeffs0"] = np.asarray(coeffs, dtype=np.float64)
        arrays["paulis0"] = np.asarray(paulis, dtype=np.uint32)
        obs = {"coeffs": "coeffs0", "paulis": "paulis0"}
    header = {"num_qubits": int(num_qubits), "density_matrix": bool(density_matrix), "ops": recs, "observable": obs}
    np.savez(path, format=np.array("fsa-circuit-1"), header=np.array(json.dumps(header)), **arrays)


def load_circuit(path):
    """(num_qubits, ops, density_matrix, coeffs or None, paulis or None) from save_circuit's file."""
    import json

    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != "fsa-circuit-1":
            raise ValueError(f"{path} is not an fsa circuit file")
        h = json.loads(str(z["header"]))
        ops = []
        for rec in h["ops"]:
            name = rec["name"]
            q = rec.get("qubits", [])
            if name == "oneTargGate":
                ops.append((name, (q[0], z[rec["matrix"]])))
            elif name == "manyCtrlOneTargGate":
                ops.append((name, (rec["controls"], q[0], z[rec["matrix"]])))
            elif name == "swapGate":
                ops.append((name, (q[0], q[1])))
            elif name == "manyTargGate":
                ops.append((name, (q, z[rec["matrix"]])))
            elif name == "pauliTensor":
                ops.append((name, (q, rec["paulis"])))
            elif name == "pauliGadget":
                ops.append((name, (q, rec["paulis"], rec["theta"])))
            elif name == "phaseGadget":
                ops.append((name, (q, rec["theta"])))
            elif name == "krausMap":
                ops.append((name, (z[rec["kraus"]], q)))
            elif name in ("twoQubitDephasing", "twoQubitDepolarising"):
                ops.append((name, (q[0], q[1], rec["prob"])))
            else:
                ops.append((name, (q[0], rec["prob"])))
        obs = h.get("observable")
        coeffs = z[obs["coeffs"]] if obs else None
        paulis = z[obs["paulis"]] if obs else None
    return h["num_qubits"], ops, h["density_matrix"], coeffs, paulis
