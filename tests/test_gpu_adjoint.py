"""GPU: golden circuits, the adjoint gradient (FP64 and FP32-bulk) and its
torch.autograd binding.

The adjoint's exp(+i theta P) gadget derivative is checked three ways:
  * against the parameter-shift gradient (f(theta + pi/4) - f(theta - pi/4),
    exact for exp(i theta P)) and the central finite differences (h = 1e-4)
    that the compiled reference produced for the 12-qubit copy of the layered
    ZZ + X circuit (tests/golden/c4_12q.npz);
  * against parameter-shift gradients from the oracle on random circuits that
    also hold non-parameterised gates (dense, controlled, swaps, tensors);
  * FP32 amplitudes against the FP64 golden values, with the error stated.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLD = ROOT / "tests" / "golden"
# FP32-bulk gradient bound (relative 2-norm vs the FP64 golden gradient)
TOL_GRAD32 = 1e-6


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


@pytest.mark.parametrize("mode", [0, 1])
def test_c1_golden_on_gpu(fsa, circ, mode):
    g = np.load(GOLD / "c1_12q.npz")
    ops = circ.config1(seed=int(g["seed"]))
    psi = fsa.StateVector(12, mode=mode)
    psi.upload(g["psi0"])
    if mode == 1:
        fsa.Circuit(ops).apply(psi)
    else:
        fsa.apply_ops(psi, ops)
    err = rel(psi.download(), g["final"])
    print(f"c1 12-qubit golden, mode {mode}: relative error {err:.3e}")
    assert err <= 1e-12


def c4_setup(fsa, circ, precision):
    g = np.load(GOLD / "c4_12q.npz")
    ops, coeffs, paulis = circ.config4(n=12)
    c = fsa.Circuit(ops)
    psi0 = fsa.StateVector(12, precision=precision)
    psi0.init_zero()
    return g, c, psi0, coeffs, paulis


def test_adjoint_fp64_matches_reference_gradients(fsa, circ):
    g, c, psi0, coeffs, paulis = c4_setup(fsa, circ, "fp64")
    expec, grads = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    e_shift, e_fd = rel(grads, g["grad_shift"]), rel(grads, g["grad_fd"])
    print(f"c4 12-qubit FP64 adjoint: <H> err {abs(expec - float(g['expec'])):.2e}, "
          f"grad vs shift {e_shift:.2e}, vs FD {e_fd:.2e}")
    assert abs(expec - float(g["expec"])) <= 1e-12
    assert e_shift <= 1e-11
    assert e_fd <= 1e-7


def test_adjoint_fp32_bulk_error_stated(fsa, circ):
    g, c, psi0, coeffs, paulis = c4_setup(fsa, circ, "fp32")
    expec, grads = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    e_grad = rel(grads, g["grad_shift"])
    e_exp = abs(expec - float(g["expec"])) / abs(float(g["expec"]))
    print(f"c4 12-qubit FP32-bulk adjoint: <H> rel err {e_exp:.2e}, grad rel err {e_grad:.2e}")
    assert e_exp <= TOL_GRAD32
    assert e_grad <= TOL_GRAD32


def random_param_circuit(rng, n, num):
    from test_gpu_parity import random_sv_op

    ops = []
    for _ in range(num):
        r = rng.random()
        q = [int(x) for x in rng.permutation(n)]
        if r < 0.35:
            k = int(rng.integers(1, 5))
            ops.append(("pauliGadget", (q[:k], [int(x) for x in rng.integers(1, 4, k)], float(rng.uniform(0, 6.3)))))
        elif r < 0.5:
            k = int(rng.integers(1, 4))
            ops.append(("phaseGadget", (q[:k], float(rng.uniform(0, 6.3)))))
        else:
            name, args = random_sv_op(rng, n)
            if name in ("oneTargGate", "manyCtrlOneTargGate", "manyTargGate"):
                u, _, vh = np.linalg.svd(args[-1])
                args = args[:-1] + (u @ vh,)
            ops.append((name, args))
    return ops


def oracle_value(orc, psi0, ops, thetas, coeffs, paulis):
    psi = psi0.copy()
    k = 0
    for name, args in ops:
        if name in ("pauliGadget", "phaseGadget"):
            args = args[:-1] + (thetas[k],)
            k += 1
        getattr(orc, "sv_" + name)(psi, *args)
    return orc.sv_innerProduct(psi, orc.sv_applyPauliSum(psi, coeffs, paulis)).real


@pytest.mark.parametrize("seed", range(4))
def test_adjoint_random_circuits_vs_oracle_shift(fsa, orc, seed):
    rng = np.random.default_rng(300 + seed)
    n = 9 if seed % 2 == 0 else 13
    ops = random_param_circuit(rng, n, 60)
    T = 12
    paulis = np.zeros((T, n), dtype=np.uint32)
    for t in range(T):
        w = int(rng.integers(1, 5))
        paulis[t, rng.choice(n, w, replace=False)] = rng.integers(1, 4, w)
    coeffs = rng.normal(size=T)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0h = v / np.linalg.norm(v)
    thetas = np.array([a[-1] for nm, a in ops if nm in ("pauliGadget", "phaseGadget")])
    shift = np.empty_like(thetas)
    for k in range(len(thetas)):
        tp, tm = thetas.copy(), thetas.copy()
        tp[k] += np.pi / 4
        tm[k] -= np.pi / 4
        shift[k] = oracle_value(orc, psi0h, ops, tp, coeffs, paulis) - oracle_value(orc, psi0h, ops, tm, coeffs, paulis)
    f0 = oracle_value(orc, psi0h, ops, thetas, coeffs, paulis)
    c = fsa.Circuit(ops)
    psi0 = fsa.StateVector(n)
    psi0.upload(psi0h)
    expec, grads = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    assert abs(expec - f0) <= 1e-12 * max(1, abs(f0))
    assert rel(grads, shift) <= 1e-11, (n, rel(grads, shift))
    # psi0 is an input only
    assert rel(psi0.download(), psi0h) == 0


def test_adjoint_reuses_its_work_states(fsa, orc):
    """The circuit keeps the adjoint's work states for the next call: calls on
    states of another size or precision rebuild them, repeated calls give the
    same bits, and work queued on the input state just before a call (an eager
    gate) is ordered before the sweep reads it."""
    rng = np.random.default_rng(41)
    n = 10
    ops = random_param_circuit(rng, n, 40)
    T = 6
    paulis = np.zeros((T, n), dtype=np.uint32)
    for t in range(T):
        paulis[t, rng.choice(n, 2, replace=False)] = rng.integers(1, 4, 2)
    coeffs = rng.normal(size=T)
    c = fsa.Circuit(ops)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    first = None
    for prec, size in (("fp64", n), ("fp32", n), ("fp64", n + 2), ("fp64", n), ("fp64", n)):
        psi0 = fsa.StateVector(size, precision=prec, mode=0)
        w = np.zeros(1 << size, dtype=np.complex128)
        w[:1 << n] = v
        psi0.upload(w)
        e, g = fsa.adjoint_gradient(c, psi0, coeffs, np.hstack([paulis, np.zeros((T, size - n), np.uint32)]))
        if prec == "fp64" and size == n:
            if first is None:
                first = (e, g.copy())
            else:
                assert e == first[0] and np.array_equal(g, first[1])
        psi0.close()
    # a gate queued on psi0 right before the call is applied before the sweep reads psi0
    psi0 = fsa.StateVector(n, mode=0)
    psi0.upload(v)
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    fsa.distributed_statevector_oneTargGate(psi0, 3, h)
    e, g = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    ref = v.copy()
    orc.sv_oneTargGate(ref, 3, h)
    thetas = np.array([a[-1] for nm, a in ops if nm in ("pauliGadget", "phaseGadget")])
    assert abs(e - oracle_value(orc, ref, ops, thetas, coeffs, paulis)) <= 1e-12


@pytest.mark.parametrize("n", [9, 12])
def test_adjoint_gadget_runs_vs_oracle_shift(fsa, orc, n):
    # the reverse sweep takes runs of consecutive diagonal gadgets (phase
    # gadgets, Z-only Pauli gadgets) up to eight per pass, and runs of one-qubit
    # X/Y gadgets on distinct qubits two or three per pass; runs end at other
    # ops, at a repeated qubit and at the run limits
    rng = np.random.default_rng(410 + n)
    u = lambda: np.linalg.svd(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    ang = lambda: float(rng.uniform(0, 6.3))
    ops = []
    for layer in range(3):
        q = [int(x) for x in rng.permutation(n)]
        for i in range(n - 1):                                         # a run of n - 1 diagonal gadgets
            if i % 4 == 3:
                ops.append(("pauliGadget", ([q[i], q[i + 1]], [3, 3], ang())))
            else:
                ops.append(("phaseGadget", ([q[i], q[i + 1]], ang())))
        ops.append(("oneTargGate", (q[0], u())))
        for grp in (q[:2], q[2:5], q[5:7]):                          # X/Y runs of 2, 3 and 2
            for x in grp:
                ops.append(("pauliGadget", ([x], [int(rng.integers(1, 3))], ang())))
            ops.append(("oneTargGate", (grp[0], u())))
        ops.append(("pauliGadget", ([q[7]], [1], ang())))              # repeated qubit ends a run
        ops.append(("pauliGadget", ([q[7]], [2], ang())))
        ops.append(("pauliGadget", ([q[8]], [2], ang())))
        for x in q[:5]:                                                # five in a row: runs of 3 and 2
            ops.append(("pauliGadget", ([x], [1], ang())))
    T = 8
    paulis = np.zeros((T, n), dtype=np.uint32)
    for t in range(T):
        w = int(rng.integers(1, 4))
        paulis[t, rng.choice(n, w, replace=False)] = rng.integers(1, 4, w)
    coeffs = rng.normal(size=T)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0h = v / np.linalg.norm(v)
    thetas = np.array([a[-1] for nm, a in ops if nm in ("pauliGadget", "phaseGadget")])
    shift = np.empty_like(thetas)
    for k in range(len(thetas)):
        tp, tm = thetas.copy(), thetas.copy()
        tp[k] += np.pi / 4
        tm[k] -= np.pi / 4
        shift[k] = oracle_value(orc, psi0h, ops, tp, coeffs, paulis) - oracle_value(orc, psi0h, ops, tm, coeffs, paulis)
    f0 = oracle_value(orc, psi0h, ops, thetas, coeffs, paulis)
    c = fsa.Circuit(ops)
    psi0 = fsa.StateVector(n)
    psi0.upload(psi0h)
    expec, grads = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    assert abs(expec - f0) <= 1e-12 * max(1, abs(f0))
    assert rel(grads, shift) <= 1e-11, (n, rel(grads, shift))
    psi32 = fsa.StateVector(n, precision="fp32")
    psi32.upload(psi0h)
    e32, g32 = fsa.adjoint_gradient(c, psi32, coeffs, paulis)
    print(f"gadget runs n={n}: fp64 rel {rel(grads, shift):.2e}, fp32 rel {rel(g32, shift):.2e}")
    assert rel(g32, shift) <= TOL_GRAD32


def test_autograd_binding(fsa, circ):
    import torch

    from fsa_amd.autograd import expectation

    g, c, psi0, coeffs, paulis = c4_setup(fsa, circ, "fp64")
    thetas = torch.tensor(g["thetas"], dtype=torch.float64, requires_grad=True)
    e = expectation(thetas, c, psi0, coeffs, paulis)
    (2.0 * e).backward()
    assert abs(e.item() - float(g["expec"])) <= 1e-12
    assert rel(thetas.grad.numpy(), 2.0 * g["grad_shift"]) <= 1e-11
    # new angles flow through set_params: a shifted angle changes <H> as the shift rule says
    k = 7
    tp = thetas.detach().clone()
    tp[k] += np.pi / 4
    tm = thetas.detach().clone()
    tm[k] -= np.pi / 4
    d = expectation(tp, c, psi0, coeffs, paulis).item() - expectation(tm, c, psi0, coeffs, paulis).item()
    assert abs(d - g["grad_shift"][k]) <= 1e-11 * max(1, abs(g["grad_shift"][k]))
