"""Closed forms and invariants that pin the oracle independently of any model.

Each check is chosen so a plausible slip (wrong sign of theta, a dropped
i^#Y, wrong column/row bit, a lost conj) fails it.
"""
import numpy as np
import pytest

import dense_model as dm


def rs(rng, n):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def test_gadget_at_half_pi_is_i_times_tensor(orc):
    rng = np.random.default_rng(1)
    psi = rs(rng, 5)
    a, b = psi.copy(), psi.copy()
    orc.sv_pauliGadget(a, [0, 3, 2], [1, 2, 3], np.pi / 2)
    orc.sv_pauliTensor(b, [0, 3, 2], [1, 2, 3])
    assert np.allclose(a, 1j * b, atol=1e-14)


def test_gadget_sign_convention_exp_plus_i_theta(orc):
    # exp(+i theta X)|0> = cos|0> + i sin|1>
    psi = np.zeros(2, complex); psi[0] = 1
    orc.sv_pauliGadget(psi, [0], [1], 0.3)
    assert np.allclose(psi, [np.cos(0.3), 1j * np.sin(0.3)], atol=1e-15)
    # exp(+i theta Y)|0> = cos|0> - sin|1>   (Y|0> = i|1>)
    psi = np.zeros(2, complex); psi[0] = 1
    orc.sv_pauliGadget(psi, [0], [2], 0.3)
    assert np.allclose(psi, [np.cos(0.3), -np.sin(0.3)], atol=1e-15)


def test_phase_gadget_parity_factors(orc):
    psi = np.ones(8, complex)
    orc.sv_phaseGadget(psi, [0, 2], 0.7)
    for j in range(8):
        par = ((j & 1) ^ ((j >> 2) & 1))
        assert np.isclose(psi[j], np.exp(-1j * 0.7) if par else np.exp(1j * 0.7))


def test_swap_is_involution_and_permutation(orc):
    rng = np.random.default_rng(2)
    psi = rs(rng, 6)
    a = psi.copy()
    orc.sv_swapGate(a, 1, 4)
    assert a[1 << 1] == psi[1 << 4] and a[1 << 4] == psi[1 << 1]
    orc.sv_swapGate(a, 4, 1)
    assert np.array_equal(a, psi)


def test_unitary_then_dagger_is_identity(orc):
    rng = np.random.default_rng(3)
    psi = rs(rng, 7)
    U = dm.haar_unitary(8, rng)
    a = psi.copy()
    orc.sv_manyTargGate(a, [5, 0, 3], U)
    assert abs(np.linalg.norm(a) - 1) < 1e-13
    orc.sv_manyTargGate(a, [5, 0, 3], U.conj().T)
    assert np.allclose(a, psi, atol=1e-13)


def test_pauli_sum_and_inner_product(orc):
    rng = np.random.default_rng(4)
    n = 5
    psi = rs(rng, n)
    codes = rng.integers(0, 4, (3, n))
    coeffs = rng.normal(size=3)
    H = sum(c * dm.pauli_kron(list(row)) for c, row in zip(coeffs, codes))
    lam = orc.sv_applyPauliSum(psi, coeffs, codes)
    assert np.allclose(lam, H @ psi, atol=1e-13)
    e = orc.sv_innerProduct(psi, lam)
    assert abs(e - np.vdot(psi, H @ psi)) < 1e-13 and abs(e.imag) < 1e-13
    t = [1, 4]
    m = orc.sv_pauliMatrixElement(psi, lam, t, [2, 1])
    P = dm.full_operator(n, t, dm.pauli_kron([2, 1]))
    assert abs(m - np.vdot(psi, P @ lam)) < 1e-13


def rand_mixed(rng, N, k=3):
    D = 1 << N
    R = np.zeros((D, D), complex)
    w = rng.dirichlet(np.ones(k))
    for wi in w:
        g = rs(rng, N)
        R += wi * np.outer(g, g.conj())
    return dm.rho_to_vec(R)


@pytest.mark.parametrize("name,nargs", [("dm_oneQubitDephasing", 1), ("dm_twoQubitDephasing", 2),
                                        ("dm_oneQubitDepolarising", 1), ("dm_twoQubitDepolarising", 2),
                                        ("dm_damping", 1)])
def test_channels_preserve_trace_and_hermiticity(orc, name, nargs):
    rng = np.random.default_rng(5)
    N = 4
    v = rand_mixed(rng, N)
    qs = [2, 0][:nargs]
    getattr(orc, name)(v, *qs, 0.37)
    R = dm.vec_to_rho(v)
    assert abs(np.trace(R) - 1) < 1e-13
    assert np.allclose(R, R.conj().T, atol=1e-14)
    assert np.linalg.eigvalsh(R).min() > -1e-12


def test_damping_full_probability_resets_qubit(orc):
    rng = np.random.default_rng(6)
    N = 3
    v = rand_mixed(rng, N)
    orc.dm_damping(v, 1, 1.0)
    R = dm.vec_to_rho(v)
    idx = np.arange(1 << N)
    assert np.allclose(R[(idx >> 1) & 1 == 1, :], 0, atol=1e-15)


def test_full_depolarising_gives_maximally_mixed_block(orc):
    rng = np.random.default_rng(7)
    N = 2
    v = rand_mixed(rng, N)
    orc.dm_oneQubitDepolarising(v, 0, 0.75)
    orc.dm_oneQubitDepolarising(v, 1, 0.75)
    assert np.allclose(dm.vec_to_rho(v), np.eye(4) / 4, atol=1e-14)
    v = rand_mixed(rng, N)
    orc.dm_twoQubitDepolarising(v, 0, 1, 15 / 16)
    assert np.allclose(dm.vec_to_rho(v), np.eye(4) / 4, atol=1e-14)


def test_single_z_expectation_is_signed_diagonal(orc):
    rng = np.random.default_rng(8)
    N = 4
    v = rand_mixed(rng, N)
    R = dm.vec_to_rho(v)
    codes = [0, 0, 3, 0]
    got = orc.dm_expecPauliString(v, [1.0], codes)
    d = np.real(np.diag(R))
    want = sum(d[j] * (1 - 2 * ((j >> 2) & 1)) for j in range(16))
    assert abs(got - want) < 1e-14


def test_partial_trace_preserves_trace(orc):
    rng = np.random.default_rng(9)
    v = rand_mixed(rng, 4)
    out = orc.dm_partialTrace(v, [3, 1])
    assert abs(np.trace(dm.vec_to_rho(out)) - 1) < 1e-14


def test_gaussian_init_is_normalised_and_seeded(orc):
    a = orc.init_gaussian(10, 123)
    b = orc.init_gaussian(10, 123)
    c = orc.init_gaussian(10, 124)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(np.linalg.norm(a) - 1) < 1e-14
    # roughly N(0,1) before normalisation: real and imaginary variances agree
    assert 0.8 < np.var(a.real) / np.var(a.imag) < 1.25
