"""GPU consumer of the on-disk circuit format (circuits.save_circuit /
load_circuit): circuits written to a file, read back and run through libfsa.

  * the 12-qubit 200-gate statevector circuit, loaded from its file and run
    op by op and as a recorded Circuit, lands on the amplitudes of the golden
    fixture written from the compiled reference (tests/golden/c1_12q.npz);
  * a slice of the density-matrix sequence with its observable, loaded from
    its file, matches the oracle run on the in-memory op list (state,
    expecPauliString from the stored observable, partialTrace);
  * the 24-qubit gadget circuit with its observable, loaded from its file,
    gives the same adjoint gradient as the in-memory circuit.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLD = ROOT / "tests" / "golden"


@pytest.fixture(scope="module")
def fsa():
    return ge.load_package()


@pytest.fixture(scope="module")
def circ():
    spec = importlib.util.spec_from_file_location("fsa_amd_circuits",
                                                  ROOT / "distributed-full-state-algorithms_amd" / "circuits.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("mode", [1, 0], ids=["fused", "eager"])
def test_c1_file_matches_golden(fsa, circ, tmp_path, mode):
    g = np.load(GOLD / "c1_12q.npz")
    path = tmp_path / "c1.npz"
    circ.save_circuit(path, 12, circ.config1(seed=int(g["seed"])))
    n, ops, dm, coeffs, paulis = circ.load_circuit(path)
    assert n == 12 and not dm and coeffs is None and len(ops) == int(g["num_gates"])
    psi = fsa.StateVector(n, mode=mode)
    psi.upload(g["psi0"])
    for name, args in ops:
        getattr(fsa, "distributed_statevector_" + name)(psi, *args)
    err = rel(psi.download(), g["final"])
    print(f"c1 from file, mode {mode}: rel err vs golden {err:.2e}")
    assert err <= 1e-12, err


def test_c1_file_as_recorded_circuit(fsa, circ, tmp_path):
    g = np.load(GOLD / "c1_12q.npz")
    path = tmp_path / "c1.npz"
    circ.save_circuit(path, 12, circ.config1(seed=int(g["seed"])))
    n, ops, _, _, _ = circ.load_circuit(path)
    c = fsa.Circuit(ops)
    psi = fsa.StateVector(n)
    psi.upload(g["psi0"])
    c.apply(psi)
    assert rel(psi.download(), g["final"]) <= 1e-12


def test_dm_file_with_observable_vs_oracle(fsa, orc, circ, tmp_path):
    N = 9
    ops, coeffs, paulis, tt = circ.config3(N=N, num_ops=120)
    path = tmp_path / "c3.npz"
    circ.save_circuit(path, N, ops, density_matrix=True, coeffs=coeffs, paulis=paulis)
    n, lops, dm, lc, lp = circ.load_circuit(path)
    assert n == N and dm and np.array_equal(lc, coeffs) and np.array_equal(lp, paulis)
    v = circ.random_mixed_state(N, 38)
    ref = v.copy()
    for name, args in ops:                    # the oracle runs the in-memory list
        getattr(orc, "dm_" + name)(ref, *args)
    rho = fsa.DensityMatrix(N)
    rho.upload(v)
    for name, args in lops:                   # the GPU runs what came back from the file
        getattr(fsa, "distributed_densitymatrix_" + name)(rho, *args)
    assert rel(rho.download(), ref) <= 1e-12
    e = fsa.distributed_densitymatrix_expecPauliString(rho, lc, lp)
    assert abs(e - orc.dm_expecPauliString(ref, coeffs, paulis)) <= 1e-12 * max(1.0, np.sum(np.abs(coeffs)))
    pt = fsa.distributed_densitymatrix_partialTrace(rho, tt)
    assert rel(pt.download(), orc.dm_partialTrace(ref, tt)) <= 1e-12


def test_gadget_circuit_file_adjoint(fsa, circ, tmp_path):
    n = 14
    ops, coeffs, paulis = circ.config4(n=n, layers=4)
    path = tmp_path / "c4.npz"
    circ.save_circuit(path, n, ops, coeffs=coeffs, paulis=paulis)
    _, lops, _, lc, lp = circ.load_circuit(path)
    res = []
    for o, c, p in ((ops, coeffs, paulis), (lops, lc, lp)):
        psi0 = fsa.StateVector(n)
        psi0.init_zero()
        res.append(fsa.adjoint_gradient(fsa.Circuit(o), psi0, c, p))
    (e0, g0), (e1, g1) = res
    assert e0 == e1 and np.array_equal(g0, g1)      # same ops, same kernels: bit-identical
