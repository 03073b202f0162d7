"""Pin the CPU oracle against the dense brute-force model.

Mirrors the reference's randomised property tests (tests/tests_statevector.hpp,
tests/tests_densitymatrix.hpp): random state, random op, compare with the dense
operator.  Random non-unitary Gaussian gate matrices as in the reference
(test_utilities.hpp:135-143), so a transposed or conjugated operand fails.
The all-Z Pauli strings the reference's tests exclude (ensureNotAllPauliZ,
test_utilities.hpp:218-224) are included here: they pin the D2 contract.
"""
import numpy as np
import pytest

import dense_model as dm

TOL = 1e-12
SV_N = 6     # tests_statevector.hpp:13
DM_N = 4     # one fewer than tests_densitymatrix.hpp:151 to keep the dense model quick
TRIALS = 40


def rand_state(rng, n):
    return rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)


def rand_mat(rng, d):
    return rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))


def close(a, b):
    return np.linalg.norm(a - b) <= TOL * max(1.0, np.linalg.norm(b))


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_one_targ(orc, seed):
    rng = np.random.default_rng(seed)
    psi = rand_state(rng, SV_N)
    t = int(rng.integers(SV_N))
    g = rand_mat(rng, 2)
    ref = dm.apply_to_vector(psi, [t], g)
    orc.sv_oneTargGate(psi, t, g)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_many_ctrl_one_targ(orc, seed):
    rng = np.random.default_rng(100 + seed)
    psi = rand_state(rng, SV_N)
    qs = rng.permutation(SV_N)
    nc = int(rng.integers(1, SV_N))
    t, ctrls = int(qs[0]), [int(q) for q in qs[1:1 + nc]]
    g = rand_mat(rng, 2)
    ref = dm.full_operator(SV_N, [t], g, ctrls) @ psi
    orc.sv_manyCtrlOneTargGate(psi, ctrls, t, g)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_swap(orc, seed):
    rng = np.random.default_rng(200 + seed)
    psi = rand_state(rng, SV_N)
    a, b = (int(x) for x in rng.choice(SV_N, 2, replace=False))
    ref = dm.apply_to_vector(psi, [a, b], dm.swap_matrix())
    orc.sv_swapGate(psi, a, b)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_many_targ(orc, seed):
    rng = np.random.default_rng(300 + seed)
    psi = rand_state(rng, SV_N)
    k = int(rng.integers(1, SV_N + 1))
    targs = [int(q) for q in rng.choice(SV_N, k, replace=False)]   # unsorted on purpose
    g = rand_mat(rng, 1 << k)
    ref = dm.apply_to_vector(psi, targs, g)
    orc.sv_manyTargGate(psi, targs, g)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_pauli_tensor(orc, seed):
    rng = np.random.default_rng(400 + seed)
    psi = rand_state(rng, SV_N)
    k = int(rng.integers(1, SV_N + 1))
    targs = [int(q) for q in rng.choice(SV_N, k, replace=False)]
    paulis = [int(p) for p in rng.integers(1, 4, k)]
    if seed % 7 == 0:
        paulis = [3] * k                        # all-Z: D2 contract
    ref = dm.apply_to_vector(psi, targs, dm.pauli_kron(paulis))
    orc.sv_pauliTensor(psi, targs, paulis)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_pauli_gadget(orc, seed):
    rng = np.random.default_rng(500 + seed)
    psi = rand_state(rng, SV_N)
    k = int(rng.integers(1, SV_N + 1))
    targs = [int(q) for q in rng.choice(SV_N, k, replace=False)]
    paulis = [int(p) for p in rng.integers(1, 4, k)]
    if seed % 7 == 0:
        paulis = [3] * k
    theta = float(rng.uniform(-np.pi, np.pi))
    ref = dm.apply_to_vector(psi, targs, dm.gadget(paulis, theta))
    orc.sv_pauliGadget(psi, targs, paulis, theta)
    assert close(psi, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_sv_phase_gadget(orc, seed):
    rng = np.random.default_rng(600 + seed)
    psi = rand_state(rng, SV_N)
    k = int(rng.integers(1, SV_N + 1))
    targs = [int(q) for q in rng.choice(SV_N, k, replace=False)]
    theta = float(rng.uniform(-np.pi, np.pi))
    ref = dm.apply_to_vector(psi, targs, dm.gadget([3] * k, theta))
    orc.sv_phaseGadget(psi, targs, theta)
    assert close(psi, ref)


# ---- density matrices ------------------------------------------------------

def rand_rho(rng, N):
    return rand_state(rng, 2 * N)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_many_targ(orc, seed):
    rng = np.random.default_rng(700 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, DM_N + 1))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    g = rand_mat(rng, 1 << k)
    ref = dm.dm_unitary(v, targs, g)
    orc.dm_manyTargGate(v, targs, g)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_swap(orc, seed):
    rng = np.random.default_rng(800 + seed)
    v = rand_rho(rng, DM_N)
    a, b = (int(x) for x in rng.choice(DM_N, 2, replace=False))
    ref = dm.dm_unitary(v, [a, b], dm.swap_matrix())
    orc.dm_swapGate(v, a, b)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_pauli_tensor(orc, seed):
    rng = np.random.default_rng(900 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, DM_N + 1))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    paulis = [int(p) for p in rng.integers(1, 4, k)]
    if seed % 7 == 0:
        paulis = [3] * k
    ref = dm.dm_unitary(v, targs, dm.pauli_kron(paulis))
    orc.dm_pauliTensor(v, targs, paulis)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_pauli_gadget(orc, seed):
    rng = np.random.default_rng(1000 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, DM_N + 1))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    paulis = [int(p) for p in rng.integers(1, 4, k)]
    if seed % 7 == 0:
        paulis = [3] * k
    theta = float(rng.uniform(-np.pi, np.pi))
    ref = dm.dm_unitary(v, targs, dm.gadget(paulis, theta))
    orc.dm_pauliGadget(v, targs, paulis, theta)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_phase_gadget(orc, seed):
    rng = np.random.default_rng(1100 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, DM_N + 1))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    theta = float(rng.uniform(-np.pi, np.pi))
    ref = dm.dm_unitary(v, targs, dm.gadget([3] * k, theta))
    orc.dm_phaseGadget(v, targs, theta)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_kraus_map(orc, seed):
    rng = np.random.default_rng(1200 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, 3))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    ops = [rand_mat(rng, 1 << k) for _ in range(int(rng.integers(1, 10)))]
    ref = dm.dm_kraus(v, targs, ops)
    orc.dm_krausMap(v, ops, targs)
    assert close(v, ref)


CHANNELS = [
    ("dm_oneQubitDephasing", 1, 0.5, dm.kraus_dephasing1),
    ("dm_twoQubitDephasing", 2, 0.75, dm.kraus_dephasing2),
    ("dm_oneQubitDepolarising", 1, 0.5, dm.kraus_depolarising1),
    ("dm_twoQubitDepolarising", 2, 0.75, dm.kraus_depolarising2),
    ("dm_damping", 1, 0.5, dm.kraus_damping),
]


@pytest.mark.parametrize("name,nq,pmax,kraus", CHANNELS, ids=[c[0] for c in CHANNELS])
@pytest.mark.parametrize("seed", range(TRIALS // 2))
def test_dm_channels(orc, name, nq, pmax, kraus, seed):
    """Kraus sets of tests_densitymatrix.hpp:262-351 (two-qubit depolarising with the textbook K0)."""
    rng = np.random.default_rng(1300 + seed)
    v = rand_rho(rng, DM_N)
    qs = [int(q) for q in rng.choice(DM_N, nq, replace=False)]
    p = float(rng.uniform(0, pmax))
    ref = dm.dm_kraus(v, qs, kraus(p))
    getattr(orc, name)(v, *qs, p)
    assert close(v, ref)


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_expec_pauli_string(orc, seed):
    """tests_densitymatrix.hpp:354-380 (there with WithinAbs 1e-12)."""
    rng = np.random.default_rng(1400 + seed)
    v = rand_rho(rng, DM_N)
    T = int(rng.integers(1, 30))
    paulis = rng.integers(0, 4, T * DM_N)
    coeffs = rng.uniform(-10, 10, T)
    got = orc.dm_expecPauliString(v, coeffs, paulis)
    want = dm.dm_expec(v, coeffs, paulis)
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("seed", range(TRIALS))
def test_dm_partial_trace(orc, seed):
    rng = np.random.default_rng(1500 + seed)
    v = rand_rho(rng, DM_N)
    k = int(rng.integers(1, DM_N + 1))
    targs = [int(q) for q in rng.choice(DM_N, k, replace=False)]
    got = orc.dm_partialTrace(v, targs)
    assert close(got, dm.dm_partial_trace(v, targs))


def test_dense_model_self_consistency():
    """The dense model's own building blocks against textbook identities."""
    rng = np.random.default_rng(7)
    U = dm.haar_unitary(4, rng)
    assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-13)
    # X on qubit 0 of |00> gives |01> (index 1); on qubit 1 gives index 2
    e0 = np.zeros(4, complex); e0[0] = 1
    assert np.argmax(abs(dm.apply_to_vector(e0, [0], dm.PAULI[1]))) == 1
    assert np.argmax(abs(dm.apply_to_vector(e0, [1], dm.PAULI[1]))) == 2
    # matrix bit i <-> targets[i]: CNOT with control = matrix bit 0 on targets [1, 0]
    cnot = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], complex)
    e = np.zeros(4, complex); e[2] = 1           # qubit 1 set
    assert np.argmax(abs(dm.apply_to_vector(e, [1, 0], cnot))) == 3
    # column-major density matrices: rho[r, c] = v[r + D c]
    v = np.arange(16).astype(complex)
    assert dm.vec_to_rho(v)[1, 2] == v[1 + 4 * 2]
