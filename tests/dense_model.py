"""Dense brute-force model of the gate semantics (test infrastructure).

An independent formulation of what the reference's test harness computes with
``getFullGateMatrix``/``applyGateToLocalState`` (tests/test_utilities.hpp:302-380):
operators are built as full 2^n x 2^n matrices by tensor contraction over the
target axes (numpy reshape/tensordot), not by the bit-insertion loops the oracle
and the GPU kernels use, so a wrong bit position, a transposed operand or a
dropped factor in those loops shows up as a mismatch here.  Only for small n.

Conventions: qubit q is bit q of the amplitude index; matrix row/column bit i
acts on targets[i]; a density matrix is the column-major 2N-qubit vector
v[r + 2^N c] = rho[r, c] (test_utilities.hpp:500).
"""
from __future__ import annotations

import numpy as np

PAULI = [
    np.array([[1, 0], [0, 1]], dtype=complex),
    np.array([[0, 1], [1, 0]], dtype=complex),
    np.array([[0, -1j], [1j, 0]], dtype=complex),
    np.array([[1, 0], [0, -1]], dtype=complex),
]


def apply_to_vector(psi: np.ndarray, targets, U: np.ndarray) -> np.ndarray:
    """Return U acting on `targets` of the n-qubit vector psi (tensor contraction)."""
    n = int(psi.size).bit_length() - 1
    k = len(targets)
    T = psi.reshape([2] * n)                      # axis a <-> qubit n-1-a
    Ut = np.asarray(U, dtype=complex).reshape([2] * (2 * k))  # out bits k-1..0, in bits k-1..0
    in_axes_U = [2 * k - 1 - i for i in range(k)]            # U input axis of matrix bit i
    axes_T = [n - 1 - t for t in targets]
    R = np.tensordot(Ut, T, axes=(in_axes_U, axes_T))
    # R axes: U out bits (k-1..0), then the remaining T axes in order
    rest = [a for a in range(n) if a not in axes_T]
    out_axes = [None] * n
    for i in range(k):                              # U out axis (k-1-i) <-> matrix bit i
        out_axes[axes_T[i]] = k - 1 - i
    for j, a in enumerate(rest):
        out_axes[a] = k + j
    R = np.transpose(R, out_axes)
    return R.reshape(-1)


def full_operator(n: int, targets, U, ctrls=()) -> np.ndarray:
    """2^n x 2^n matrix of U on targets, controlled on all ctrls being 1."""
    D = 1 << n
    F = np.zeros((D, D), dtype=complex)
    eye = np.eye(D, dtype=complex)
    for c in range(D):
        F[:, c] = apply_to_vector(eye[:, c].copy(), targets, U)
    if ctrls:
        idx = np.arange(D)
        on = np.ones(D, dtype=bool)
        for q in ctrls:
            on &= ((idx >> q) & 1) == 1
        proj = np.diag(on.astype(complex))
        F = eye + proj @ (F - eye)
    return F


def pauli_kron(paulis) -> np.ndarray:
    """Matrix of the tensor product with paulis[i] on matrix bit i."""
    M = np.array([[1]], dtype=complex)
    for p in paulis:
        M = np.kron(PAULI[p], M)
    return M


def gadget(paulis, theta) -> np.ndarray:
    P = pauli_kron(paulis)
    return np.cos(theta) * np.eye(P.shape[0]) + 1j * np.sin(theta) * P


def swap_matrix() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


# ---- density matrices --------------------------------------------------------

def vec_to_rho(v: np.ndarray) -> np.ndarray:
    D = int(round(np.sqrt(v.size)))
    return v.reshape(D, D).T.copy()


def rho_to_vec(R: np.ndarray) -> np.ndarray:
    return R.T.reshape(-1).copy()


def dm_unitary(v: np.ndarray, targets, U, ctrls=()) -> np.ndarray:
    R = vec_to_rho(v)
    N = int(R.shape[0]).bit_length() - 1
    F = full_operator(N, targets, U, ctrls)
    return rho_to_vec(F @ R @ F.conj().T)


def dm_kraus(v: np.ndarray, targets, kraus_ops) -> np.ndarray:
    R = vec_to_rho(v)
    N = int(R.shape[0]).bit_length() - 1
    out = np.zeros_like(R)
    for K in kraus_ops:
        F = full_operator(N, targets, K)
        out += F @ R @ F.conj().T
    return rho_to_vec(out)


def kraus_dephasing1(p):
    return [np.sqrt(1 - p) * PAULI[0], np.sqrt(p) * PAULI[3]]


def kraus_dephasing2(p):
    I, Z = PAULI[0], PAULI[3]
    return [np.sqrt(1 - p) * np.kron(I, I), np.sqrt(p / 3) * np.kron(I, Z),
            np.sqrt(p / 3) * np.kron(Z, I), np.sqrt(p / 3) * np.kron(Z, Z)]


def kraus_depolarising1(p):
    return [np.sqrt(1 - p) * PAULI[0]] + [np.sqrt(p / 3) * PAULI[i] for i in (1, 2, 3)]


def kraus_depolarising2(p):
    """Textbook channel: K0 = sqrt(1-p) II, and sqrt(p/15) P for the 15 P != II."""
    ops = []
    for a in range(4):
        for b in range(4):
            ops.append(np.sqrt(p / 15) * np.kron(PAULI[a], PAULI[b]))
    ops[0] = np.sqrt(1 - p) * np.eye(4, dtype=complex)
    return ops


def kraus_damping(p):
    return [np.array([[1, 0], [0, np.sqrt(1 - p)]], dtype=complex),
            np.array([[0, np.sqrt(p)], [0, 0]], dtype=complex)]


def dm_expec(v: np.ndarray, coeffs, all_paulis) -> complex:
    R = vec_to_rho(v)
    N = int(R.shape[0]).bit_length() - 1
    H = np.zeros_like(R)
    codes = np.asarray(all_paulis).reshape(len(coeffs), N)
    for c, row in zip(coeffs, codes):
        H += c * pauli_kron(list(row))
    return complex(np.trace(H @ R))


def dm_partial_trace(v: np.ndarray, targets) -> np.ndarray:
    R = vec_to_rho(v)
    N = int(R.shape[0]).bit_length() - 1
    T = R.reshape([2] * (2 * N))        # row axes: qubit N-1-a ; col axes: qubit 2N-1-a
    keep = [q for q in range(N) if q not in targets]
    # einsum: traced qubits share a label between their row and column axes
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    row_l, col_l = {}, {}
    for q in range(N):
        row_l[q] = next(letters)
        col_l[q] = row_l[q] if q in targets else next(letters)
    src = "".join(row_l[N - 1 - a] for a in range(N)) + "".join(col_l[N - 1 - a] for a in range(N))
    dst = "".join(row_l[q] for q in reversed(keep)) + "".join(col_l[q] for q in reversed(keep))
    out = np.einsum(src + "->" + dst, T)
    d = 1 << len(keep)
    return rho_to_vec(out.reshape(d, d))


def haar_unitary(d: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random U(d): QR of a complex Ginibre matrix with the phase of diag(R) removed."""
    Z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    Q, R = np.linalg.qr(Z)
    ph = np.diag(R) / np.abs(np.diag(R))
    return Q * ph[None, :]
