"""Dense gates of 6-12 targets as the MFMA GEMM of kernels_gemm.hip, against
the oracle (local_statevector.hpp:72-99; the Kraus map's superoperator,
distributed_densitymatrix.hpp:79-89 and misc.hpp:58-81).

Gates of 6-8 targets run in place (one workgroup holds every row of its
column block); 9-12 targets split the rows over workgroups and go through the
scratch area, which FSA_GEMM_SCRATCH_KB shrinks here so that a gate takes
several batches.  4 and 5 targets (the per-group kernel, or tile passes in
fused mode) are the neighbours on either side of the switch."""
import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
# FP32 storage, unitary gates: each output sums 2^k products of FP32 inputs
TOL32 = 1e-5


@pytest.fixture(scope="module")
def fsa():
    return ge.load_package()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def unitary(rng, d):
    u, _, vh = np.linalg.svd(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return u @ vh


def state(rng, n):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


CASES = [(4, 14), (5, 16), (6, 12), (6, 20), (7, 18), (8, 20), (9, 20), (10, 18), (11, 17), (12, 18)]


@pytest.mark.parametrize("k,n", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_many_targ_gate_gemm(fsa, orc, k, n, mode, monkeypatch):
    print(f"gemm k={k} n={n} mode={mode}", flush=True)
    if k >= 9:
        monkeypatch.setenv("FSA_GEMM_SCRATCH_KB", "4096")
    rng = np.random.default_rng(100 * k + n + mode)
    perm = [int(x) for x in rng.permutation(n)]
    # low qubits among the targets (column blocks made of short runs) and not
    targets = [[0, 1] + [q for q in perm if q > 1][:k - 2], [q for q in perm if q >= 3][:k]]
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        psi = fsa.StateVector(n, precision=prec, mode=mode)
        v = state(rng, n)
        psi.upload(v)
        ref = v.copy()
        for t in targets:
            u = unitary(rng, 1 << k)
            fsa.distributed_statevector_manyTargGate(psi, t, u)
            orc.sv_manyTargGate(ref, t, u)
        err = rel(psi.download(), ref)
        print(f"  {prec} rel err {err:.3e}")
        assert err <= tol, (k, n, mode, prec, err)


def test_gemm_matches_per_group_kernels(fsa, monkeypatch):
    """The GEMM against the per-group kernel (FSA_GEMM_MINK=13) on the same
    non-unitary gate, FP64, eager: two independent GPU forms."""
    rng = np.random.default_rng(7)
    n, outs = 20, []
    targets = [3, 17, 0, 9, 12, 5, 1]
    g = rng.normal(size=(128, 128)) + 1j * rng.normal(size=(128, 128))
    v = state(rng, n)
    for mink in ("6", "13"):
        monkeypatch.setenv("FSA_GEMM_MINK", mink)
        psi = fsa.StateVector(n, mode=0)
        psi.upload(v)
        fsa.distributed_statevector_manyTargGate(psi, targets, g)
        outs.append(psi.download())
    assert rel(outs[0], outs[1]) <= TOL64


@pytest.mark.parametrize("k,N", [(3, 8), (3, 10), (4, 9), (5, 8)])
def test_kraus_map_gemm(fsa, orc, k, N):
    """krausMap on k targets: a 2k-target superoperator on the 2N-qubit vector."""
    print(f"kraus k={k} N={N}", flush=True)
    rng = np.random.default_rng(300 + 10 * k + N)
    q = [int(x) for x in rng.permutation(N)[:k]]
    ops = np.stack([(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))) * 0.3
                    for _ in range(3)])
    for mode in (0, 1):
        rho = fsa.DensityMatrix(N, mode=mode)
        v = state(rng, 2 * N)
        rho.upload(v)
        ref = v.copy()
        fsa.distributed_densitymatrix_krausMap(rho, ops, q)
        orc.dm_krausMap(ref, ops, q)
        fsa.distributed_densitymatrix_oneQubitDephasing(rho, q[0], 0.2)
        orc.dm_oneQubitDephasing(ref, q[0], 0.2)
        err = rel(rho.download(), ref)
        print(f"  mode {mode} rel err {err:.3e}")
        assert err <= TOL64, (k, N, mode, err)
