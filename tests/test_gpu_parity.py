"""GPU parity: every operation through the C ABI against the CPU oracle.

Mirrors the reference's randomised property tests (tests/tests_statevector.hpp,
tests/tests_densitymatrix.hpp) with the oracle as the checker instead of the
dense matrix (the oracle itself is pinned against the dense model and the
compiled reference in test_oracle_*.py).  Both execution modes are covered:
eager (one streaming kernel per op) and fused (ops collected and applied in
LDS-tile passes).  FP64 tolerance: relative 2-norm error <= 1e-12.
"""
import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-5
MODES = [0, 1]          # FSA_EXEC_EAGER, FSA_EXEC_FUSED


@pytest.fixture(scope="module")
def fsa():
    return ge.load_package()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def rs(rng, n):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def rm(rng, d):
    return rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))


def random_sv_op(rng, n, kinds=("one", "ctrl", "swap", "many", "tensor", "gadget", "phase")):
    kind = kinds[int(rng.integers(len(kinds)))]
    q = [int(x) for x in rng.permutation(n)]
    if kind == "one":
        return ("oneTargGate", (q[0], rm(rng, 2)))
    if kind == "ctrl" and n >= 2:
        nc = int(rng.integers(1, min(n - 1, 4) + 1))
        return ("manyCtrlOneTargGate", (q[1:1 + nc], q[0], rm(rng, 2)))
    if kind == "swap" and n >= 2:
        return ("swapGate", (q[0], q[1]))
    if kind == "many":
        k = int(rng.integers(1, min(n, 5) + 1))
        return ("manyTargGate", (q[:k], rm(rng, 1 << k)))
    if kind in ("tensor", "gadget"):
        k = int(rng.integers(1, min(n, 6) + 1))
        p = [int(x) for x in rng.integers(1, 4, k)]
        if rng.random() < 0.15:
            p = [3] * k
        if kind == "tensor":
            return ("pauliTensor", (q[:k], p))
        return ("pauliGadget", (q[:k], p, float(rng.uniform(-np.pi, np.pi))))
    k = int(rng.integers(1, min(n, 6) + 1))
    return ("phaseGadget", (q[:k], float(rng.uniform(-np.pi, np.pi))))


def apply_sv(fsa, orc, psi, ref, name, args):
    getattr(fsa, "distributed_statevector_" + name)(psi, *args)
    getattr(orc, "sv_" + name)(ref, *args)


@pytest.mark.parametrize("mode", MODES, ids=["eager", "fused"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 12, 13, 17])
def test_sv_single_ops(fsa, orc, mode, n):
    rng = np.random.default_rng(1000 * n + mode)
    trials = 30 if n <= 13 else 10
    psi = fsa.StateVector(n, mode=mode)
    for t in range(trials):
        name, args = random_sv_op(rng, n)
        v = rs(rng, n)
        psi.upload(v)
        ref = v.copy()
        apply_sv(fsa, orc, psi, ref, name, args)
        err = rel(psi.download(), ref)
        assert err <= TOL64, (n, t, name, args[:-1] if name != "swapGate" else args, err)


@pytest.mark.parametrize("n", [6, 7, 12])
def test_sv_one_target_every_qubit_eager(fsa, orc, n):
    # every target, FP64 and FP32: targets inside a 128-byte line take the
    # lane-exchange kernel (uncontrolled), the others the pair kernel
    rng = np.random.default_rng(77 + n)
    for prec, tol in (("fp64", TOL64), ("fp32", 1e-6)):
        psi = fsa.StateVector(n, mode=0, precision=prec)
        for t in range(n):
            u = np.linalg.svd(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
            v = rs(rng, n)
            psi.upload(v)
            ref = v.copy()
            apply_sv(fsa, orc, psi, ref, "oneTargGate", (t, u))
            err = rel(psi.download(), ref)
            assert err <= tol, (prec, n, t, err)


TUNINGS = {"default": {}, "valu-dense": {"FSA_MFMA": "0"}, "sweeps": {"FSA_SWEEPS": "1"},
           "register-tiles": {"FSA_RTILE": "1"}, "lds-tiles": {"FSA_RTILE": "0"},
           "sweeps-valu": {"FSA_SWEEPS": "1", "FSA_MFMA": "0"}, "tile13": {"FSA_TILE_BITS": "13"},
           "tile13-low5": {"FSA_TILE_BITS": "13", "FSA_TILE_LOW": "5"},
           # the plain greedy scheduler; 2^11 tiles (four workgroups per CU, two-pair
           # batches); the fixed two workgroups per CU
           "greedy-schedule": {"FSA_SCHED_SEEDS": "0"}, "tile11": {"FSA_TILE_BITS": "11"},
           "fixed-grid": {"FSA_TILE_OCC": "0", "FSA_TILE_BITS": "11"},
           # tiles walked in index order instead of bit-reversed order
           "natural-order": {"FSA_TILE_ORDER": "0"}}


@pytest.mark.parametrize("n", [4, 9, 12, 14, 18, 20])
@pytest.mark.parametrize("tuning", list(TUNINGS))
def test_sv_fused_circuits(fsa, orc, n, tuning, monkeypatch):
    """Long random circuits in fused mode: scheduler + tile kernel, one readback.
    Every tuning switch of the pass builder (dense gates on MFMA or VALU,
    register sweeps on or off) must give the same amplitudes."""
    for k, v in TUNINGS[tuning].items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(77 + n)
    psi = fsa.StateVector(n, mode=1)
    v = rs(rng, n)
    psi.upload(v)
    ref = v.copy()
    for _ in range(120):
        name, args = random_sv_op(rng, n)
        apply_sv(fsa, orc, psi, ref, name, args)
    assert rel(psi.download(), ref) <= TOL64


@pytest.mark.parametrize("n", [14, 17])
@pytest.mark.parametrize("tuning", ["default", "valu-dense", "tile13", "register-tiles"])
def test_dense_only_circuits(fsa, orc, n, tuning, monkeypatch):
    """Passes holding only 3-5 target dense gates (the MFMA form on full tiles of
    2^12 and 2^13 amplitudes, the VALU form otherwise), unitary matrices."""
    for k, v in TUNINGS[tuning].items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(900 + n)
    for precision in ("fp64", "fp32"):
        psi = fsa.StateVector(n, precision=precision, mode=1)
        v = rs(rng, n)
        psi.upload(v)
        ref = v.copy()
        for _ in range(30):
            k = int(rng.integers(3, 6))
            q = [int(x) for x in rng.permutation(n)[:k]]
            u, _, vh = np.linalg.svd(rm(rng, 1 << k))
            apply_sv(fsa, orc, psi, ref, "manyTargGate", (q, u @ vh))
        tol = TOL64 if precision == "fp64" else TOL32 * np.sqrt(30)
        err = rel(psi.download(), ref)
        print(f"dense-only n={n} {tuning} {precision}: rel err {err:.3e}")
        assert err <= tol, (precision, tuning)


def test_sv_gadget_limits(fsa, orc):
    """Gadget on every qubit (too wide for one tile) and controls on every other qubit."""
    n = 16
    rng = np.random.default_rng(5)
    for mode in MODES:
        psi = fsa.StateVector(n, mode=mode)
        v = rs(rng, n)
        psi.upload(v)
        ref = v.copy()
        apply_sv(fsa, orc, psi, ref, "pauliGadget", (list(range(n)), [1, 2] * (n // 2), 0.3))
        apply_sv(fsa, orc, psi, ref, "manyCtrlOneTargGate", (list(range(1, n)), 0, rm(rng, 2)))
        apply_sv(fsa, orc, psi, ref, "manyTargGate", ([15, 0, 7, 3, 9, 12, 1], rm(rng, 128)))
        assert rel(psi.download(), ref) <= TOL64


# ---- density matrices ---------------------------------------------------------

def random_dm_op(rng, N):
    q = [int(x) for x in rng.permutation(N)]
    kind = int(rng.integers(12))
    p = float(rng.uniform(0, 0.5))
    if kind == 0:
        k = int(rng.integers(1, min(N, 3) + 1))
        return ("manyTargGate", (q[:k], rm(rng, 1 << k)))
    if kind == 1 and N >= 2:
        return ("swapGate", (q[0], q[1]))
    if kind == 2:
        k = int(rng.integers(1, N + 1))
        return ("pauliTensor", (q[:k], [int(x) for x in rng.integers(1, 4, k)]))
    if kind == 3:
        k = int(rng.integers(1, N + 1))
        return ("pauliGadget", (q[:k], [int(x) for x in rng.integers(1, 4, k)], float(rng.uniform(-3, 3))))
    if kind == 4:
        k = int(rng.integers(1, N + 1))
        return ("phaseGadget", (q[:k], float(rng.uniform(-3, 3))))
    if kind == 5:
        k = int(rng.integers(1, min(N, 2) + 1))
        ops = rm(rng, 1 << k)[None, :, :] * np.ones((int(rng.integers(1, 4)), 1, 1))
        ops = ops + rng.normal(size=ops.shape) * 0.1
        return ("krausMap", (ops, q[:k]))
    if kind == 6:
        return ("oneQubitDephasing", (q[0], p))
    if kind == 7 and N >= 2:
        return ("twoQubitDephasing", (q[0], q[1], p))
    if kind == 8:
        return ("oneQubitDepolarising", (q[0], p))
    if kind == 9 and N >= 2:
        return ("twoQubitDepolarising", (q[0], q[1], p))
    return ("damping", (q[0], p))


def apply_dm(fsa, orc, rho, ref, name, args):
    getattr(fsa, "distributed_densitymatrix_" + name)(rho, *args)
    getattr(orc, "dm_" + name)(ref, *args)


@pytest.mark.parametrize("mode", MODES, ids=["eager", "fused"])
@pytest.mark.parametrize("N", [1, 2, 3, 5, 6, 7, 9])
def test_dm_single_ops(fsa, orc, mode, N):
    rng = np.random.default_rng(2000 * N + mode)
    rho = fsa.DensityMatrix(N, mode=mode)
    for t in range(30):
        name, args = random_dm_op(rng, N)
        v = rs(rng, 2 * N)
        rho.upload(v)
        ref = v.copy()
        apply_dm(fsa, orc, rho, ref, name, args)
        err = rel(rho.download(), ref)
        assert err <= TOL64, (N, t, name, err)


@pytest.mark.parametrize("N", [3, 6, 8, 10])
def test_dm_fused_sequences(fsa, orc, N):
    rng = np.random.default_rng(300 + N)
    rho = fsa.DensityMatrix(N, mode=1)
    v = rs(rng, 2 * N)
    rho.upload(v)
    ref = v.copy()
    for _ in range(60):
        name, args = random_dm_op(rng, N)
        apply_dm(fsa, orc, rho, ref, name, args)
    assert rel(rho.download(), ref) <= TOL64


@pytest.mark.parametrize("N", [3, 6, 9])
@pytest.mark.parametrize("mode", MODES, ids=["eager", "fused"])
def test_dm_fp32_against_fp64_oracle(fsa, orc, N, mode):
    """FP32 density matrices: 40 random ops (unitary dense gates, channels, Kraus
    maps on up to 3 targets, i.e. the 6-target GEMM), then expecPauliString and
    partialTrace, against the FP64 oracle."""
    rng = np.random.default_rng(900 + 10 * N + mode)
    rho = fsa.DensityMatrix(N, precision="fp32", mode=mode)
    v = rs(rng, 2 * N)
    rho.upload(v)
    ref = v.copy()
    for _ in range(40):
        name, args = random_dm_op(rng, N)
        if name == "manyTargGate":
            u, _, vh = np.linalg.svd(args[1])
            args = (args[0], u @ vh)
        if name == "krausMap" and N >= 3 and rng.random() < 0.5:
            q = [int(x) for x in rng.permutation(N)[:3]]
            ops = np.stack([unitary_like(rng, 8) / np.sqrt(2) for _ in range(2)])
            args = (ops, q)
        apply_dm(fsa, orc, rho, ref, name, args)
    err = rel(rho.download(), ref)
    print(f"fp32 dm N={N} mode={mode}: rel err {err:.3e}")
    assert err <= 2e-6, err
    T = 8
    paulis = rng.integers(0, 4, T * N)
    coeffs = rng.normal(size=T)
    e = fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)
    e_ref = orc.dm_expecPauliString(ref, coeffs, paulis)
    assert abs(e - e_ref) <= 2e-6 * max(1.0, abs(e_ref)) * np.linalg.norm(coeffs)
    k = min(2, N - 1)
    targs = [int(x) for x in rng.permutation(N)[:k]]
    out = fsa.distributed_densitymatrix_partialTrace(rho, targs)
    assert rel(out.download(), orc.dm_partialTrace(ref, targs)) <= 2e-6


def unitary_like(rng, d):
    u, _, vh = np.linalg.svd(rm(rng, d))
    return u @ vh


@pytest.mark.parametrize("N", [1, 3, 5, 8, 11])
def test_dm_expec_pauli_string(fsa, orc, N):
    rng = np.random.default_rng(400 + N)
    rho = fsa.DensityMatrix(N)
    v = rs(rng, 2 * N)
    rho.upload(v)
    T = int(rng.integers(1, 30))
    paulis = rng.integers(0, 4, T * N)
    coeffs = rng.uniform(-10, 10, T)
    got = fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)
    want = orc.dm_expecPauliString(v, coeffs, paulis) if N <= 8 else None
    if want is None:   # large N: the oracle's full loop is slow; use the sparse identity
        want = 0
        D = 1 << N
        r = np.arange(D)
        for t in range(T):
            codes = paulis[t * N:(t + 1) * N]
            mxy = sum(1 << q for q in range(N) if codes[q] in (1, 2))
            myz = sum(1 << q for q in range(N) if codes[q] in (2, 3))
            ny = int(np.sum(codes == 2))
            c = r ^ mxy
            sg = 1 - 2 * (np.vectorize(lambda x: bin(x).count("1") & 1)(c & myz))
            want += coeffs[t] * (-1j) ** ny * np.sum(sg * v[r + D * c])
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("N,k", [(1, 1), (3, 1), (3, 2), (5, 3), (5, 5), (8, 4), (10, 2)])
def test_dm_partial_trace(fsa, orc, N, k):
    rng = np.random.default_rng(500 + N * 10 + k)
    rho = fsa.DensityMatrix(N)
    v = rs(rng, 2 * N)
    rho.upload(v)
    targs = [int(x) for x in rng.choice(N, k, replace=False)]
    out = fsa.distributed_densitymatrix_partialTrace(rho, targs)
    assert out.numQubits == N - k
    assert rel(out.download(), orc.dm_partialTrace(v, targs)) <= TOL64


def test_dm_two_qubit_depolarising_as_shipped(fsa, orc):
    rng = np.random.default_rng(9)
    v = rs(rng, 8)
    fsa.set_two_qubit_depolarising_as_shipped(True)
    try:
        rho = fsa.DensityMatrix(4, mode=0)
        rho.upload(v)
        fsa.distributed_densitymatrix_twoQubitDepolarising(rho, 3, 1, 0.4)
        ref = v.copy()
        orc.dm_twoQubitDepolarising_asShipped(ref, 3, 1, 0.4)
        assert rel(rho.download(), ref) <= TOL64
    finally:
        fsa.set_two_qubit_depolarising_as_shipped(False)


# ---- precision, init, reductions ------------------------------------------------

def test_gaussian_init_matches_oracle_law(fsa, orc):
    for n in (3, 12, 21):
        psi = fsa.StateVector(n)
        psi.init_gaussian(1234)
        got = psi.download()
        assert rel(got, orc.init_gaussian(n, 1234)) <= 1e-13
        assert abs(psi.norm2() - 1) < 1e-13


@pytest.mark.parametrize("n", [6, 14, 19])
@pytest.mark.parametrize("tuning", list(TUNINGS))
def test_fp32_bulk_against_fp64_oracle(fsa, orc, n, tuning, monkeypatch):
    """FP32 amplitudes (every kernel form, dense gates included) against the
    FP64 oracle, with unitary gates so the error does not grow with the norm."""
    for k, v in TUNINGS[tuning].items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(600 + n)
    for mode in MODES:
        psi = fsa.StateVector(n, precision="fp32", mode=mode)
        v = rs(rng, n)
        psi.upload(v)
        ref = v.copy()
        for _ in range(40):
            name, args = random_sv_op(rng, n, kinds=("one", "ctrl", "swap", "gadget", "phase", "tensor", "many"))
            if name in ("oneTargGate", "manyCtrlOneTargGate", "manyTargGate"):
                g = args[-1]
                u, _, vh = np.linalg.svd(g)
                args = args[:-1] + (u @ vh,)
            apply_sv(fsa, orc, psi, ref, name, args)
        err = rel(psi.download(), ref)
        print(f"fp32 bulk n={n} {tuning} mode={mode}: rel err {err:.3e}")
        # 40 unitary FP32 gates give at most 3.7e-7 here (every tuning, n = 6-19)
        assert err <= 1.5e-6


def test_sv_expectation_and_inner_products(fsa, orc):
    rng = np.random.default_rng(8)
    n = 13
    v = rs(rng, n)
    psi = fsa.StateVector(n)
    psi.upload(v)
    codes = rng.integers(0, 4, (20, n))
    coeffs = rng.normal(size=20)
    lam_ref = orc.sv_applyPauliSum(v, coeffs, codes)
    e = fsa.statevector_expecPauliSum(psi, coeffs, codes)
    assert abs(e - np.vdot(v, lam_ref)) < 1e-12
    lam = fsa.StateVector(n)
    fsa.statevector_applyPauliSum(psi, coeffs, codes, lam)
    assert rel(lam.download(), lam_ref) <= TOL64
    ip = fsa.statevector_innerProduct(psi, lam)
    assert abs(ip - np.vdot(v, lam_ref)) < 1e-12


def test_errors_surface_as_exceptions(fsa):
    psi = fsa.StateVector(3)
    with pytest.raises(fsa.FsaError) as e:
        fsa.distributed_statevector_oneTargGate(psi, 3, np.eye(2))
    assert e.value.code == 2
    with pytest.raises(fsa.FsaError) as e:
        fsa.distributed_statevector_manyTargGate(psi, [1, 1], np.eye(4))
    assert e.value.code == 3
    with pytest.raises(fsa.FsaError) as e:
        fsa.distributed_densitymatrix_damping(psi, 0, 0.1)
    assert e.value.code == 7


def test_deterministic_reruns(fsa):
    """Same inputs twice -> bitwise identical outputs (no racy writes, fixed-order reductions)."""
    rng = np.random.default_rng(11)
    n = 16
    v = rs(rng, n)
    outs, norms = [], []
    for _ in range(2):
        psi = fsa.StateVector(n)
        psi.upload(v)
        r2 = np.random.default_rng(3)
        for _ in range(50):
            name, args = random_sv_op(r2, n)
            getattr(fsa, "distributed_statevector_" + name)(psi, *args)
        outs.append(psi.download())
        norms.append(psi.norm2())
    assert np.array_equal(outs[0], outs[1]) and norms[0] == norms[1]


@pytest.mark.parametrize("n", [12, 16])
def test_wide_dense_gate_between_fused_passes(fsa, orc, n):
    """An 8-target gate fits no tile: it runs as a standalone kernel between two
    tile passes and pushes its 1 MiB matrix while the later pass's items are
    already in the parameter ring (they must stay valid)."""
    rng = np.random.default_rng(40 + n)
    psi = fsa.StateVector(n, mode=1)
    v = rs(rng, n)
    psi.upload(v)
    ref = v.copy()
    ops = [("oneTargGate", (3, rm(rng, 2))), ("oneTargGate", (9, rm(rng, 2))),
           ("manyTargGate", ([5, 0, 7, 2, 11, 3, 9, 6], np.linalg.svd(rm(rng, 256))[0])),
           ("oneTargGate", (0, rm(rng, 2))), ("oneTargGate", (7, rm(rng, 2))),
           ("manyTargGate", ([1, 10, 4, 8, 2, 0, 11, 5, 3], np.linalg.svd(rm(rng, 512))[0])),
           ("pauliGadget", ([2, 10], [1, 2], 0.4)), ("oneTargGate", (11, rm(rng, 2)))]
    for _ in range(3):                      # several flushes reuse and outgrow the ring
        for name, args in ops:
            apply_sv(fsa, orc, psi, ref, name, args)
        psi.flush()
    err = rel(psi.download(), ref)
    assert err <= TOL64, err


def test_dm_kraus_map_on_four_targets(fsa, orc):
    # a 4-target Kraus map is an 8-target superoperator: the standalone dense path
    rng = np.random.default_rng(41)
    N = 6
    rho = fsa.DensityMatrix(N, mode=1)
    v = rs(rng, 2 * N)
    rho.upload(v)
    ref = v.copy()
    kraus = np.stack([rm(rng, 16) * 0.3 for _ in range(3)])
    apply_dm(fsa, orc, rho, ref, "oneQubitDepolarising", (1, 0.2))
    apply_dm(fsa, orc, rho, ref, "damping", (4, 0.3))
    apply_dm(fsa, orc, rho, ref, "krausMap", (kraus, [4, 0, 5, 2]))
    apply_dm(fsa, orc, rho, ref, "oneQubitDephasing", (0, 0.1))
    apply_dm(fsa, orc, rho, ref, "manyTargGate", ([3, 5], rm(rng, 4)))
    assert rel(rho.download(), ref) <= TOL64


def test_copy_is_ordered_before_later_source_work(fsa, orc):
    """copy_from queues the copy on the destination's stream; an op on the source
    right after it must not overwrite the source before the copy has read it."""
    rng = np.random.default_rng(42)
    n = 26
    src = fsa.StateVector(n, mode=0)
    src.init_gaussian(77)
    dst = fsa.StateVector(n, mode=0)
    dst.copy_from(src)
    u = np.linalg.svd(rm(rng, 2))[0]
    for t in range(6):                      # work queued on the source's stream at once
        fsa.distributed_statevector_oneTargGate(src, t + 20, u)
    want = orc.init_gaussian(n, 77)
    # the device and oracle laws agree to 1e-13 (normalisation sums in another
    # order); a copy of a half-updated source would be off by O(1)
    assert rel(dst.download(), want) <= 1e-13
    for t in range(6):
        orc.sv_oneTargGate(want, t + 20, u)
    assert rel(src.download(), want) <= TOL64


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-14), ("fp32", 3e-7)])
def test_statevector_renormalise(fsa, orc, precision, tol):
    """fsa_statevector_renormalise: the FP64 norm of the device state, then one
    scaling pass; the direction is unchanged."""
    rng = np.random.default_rng(77)
    n = 18
    v = rs(rng, n)
    psi = fsa.StateVector(n, precision=precision)
    psi.upload(3.0 * v)
    psi.renormalise()
    assert abs(psi.norm2() - 1) <= tol
    assert rel(psi.download(), v) <= (1e-15 if precision == "fp64" else 1e-6)
    # queued ops are applied first: the norm is that of the final state
    psi.set_exec_mode(1)
    fsa.distributed_statevector_oneTargGate(psi, 4, 2.0 * np.eye(2))
    psi.renormalise()
    assert abs(psi.norm2() - 1) <= tol
    zero = fsa.StateVector(8, precision=precision)
    zero.renormalise()
    assert np.all(zero.download() == 0)
    rho = fsa.DensityMatrix(3)
    assert fsa.lib().fsa_statevector_renormalise(rho.handle) == 7      # FSA_ERR_KIND
