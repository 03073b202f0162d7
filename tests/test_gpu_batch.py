"""GPU: batched circuits with adjoint gradients (one persistent kernel).

Every job of a batch must give the same <H> and gradient as the single-state
adjoint (fsa_adjoint_gradient, itself pinned against reference-computed
parameter-shift gradients and the oracle) on the same circuit, for LDS-
resident states and for states streamed through the arena, in FP64 (1e-12)
and FP32 (stated error), across configurations, graph replays and angle
updates.  A few jobs are also checked against the oracle's shift rule.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


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


def single_adjoint(fsa, n, ops, coeffs, paulis, precision="fp64"):
    c = fsa.Circuit(ops)
    psi0 = fsa.StateVector(n, precision=precision)
    psi0.init_zero()
    e, g = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
    psi0.close()
    return e, g


def small_batch(circ, seed, num, nmin, nmax, terms=6):
    return circ.config5(seed=seed, num=num, nmin=nmin, nmax=nmax, num_terms=terms)


@pytest.mark.parametrize("cfg", [None, (256, 10, 5), (512, 11, 4), (256, 12, 6)])
def test_batch_fp64_matches_single_adjoint(fsa, circ, cfg):
    jobs = small_batch(circ, 11, 24, 9, 14)
    b = fsa.Batch("fp64")
    if cfg:
        b.set_config(*cfg)
    for n, ops, coeffs, paulis in jobs:
        b.add(fsa.Circuit(ops), n, coeffs, paulis)
    e, g = b.run()
    worst = 0.0
    for j, (n, ops, coeffs, paulis) in enumerate(jobs):
        e1, g1 = single_adjoint(fsa, n, ops, coeffs, paulis)
        assert abs(e[j] - e1) <= 1e-12 * max(1, abs(e1)), (j, n)
        worst = max(worst, rel(g[j], g1))
        assert rel(g[j], g1) <= 1e-12, (j, n, rel(g[j], g1))
    print(f"batch FP64 config {b.config}: worst gradient rel err vs single adjoint {worst:.2e}")


def test_batch_replay_and_new_angles(fsa, circ):
    jobs = small_batch(circ, 12, 10, 10, 15)
    b = fsa.Batch("fp64")
    b.set_config(256, 11, 5)
    for n, ops, coeffs, paulis in jobs:
        b.add(fsa.Circuit(ops), n, coeffs, paulis)
    e0, g0 = b.run()
    for _ in range(3):                       # graph replays are bit-identical
        e, g = b.run()
        assert np.array_equal(e, e0)
        assert all(np.array_equal(x, y) for x, y in zip(g, g0))
    rng = np.random.default_rng(5)
    for j in (0, 3, 7):
        n, ops, coeffs, paulis = jobs[j]
        th = rng.uniform(0, 2 * np.pi, len(g0[j]))
        b.set_params(j, th)
        c = fsa.Circuit(ops)
        c.set_params(th)
        psi0 = fsa.StateVector(n)
        psi0.init_zero()
        e1, g1 = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
        jobs[j] = (n, ops, coeffs, paulis, e1, g1)
    e, g = b.run()
    for j in (0, 3, 7):
        _, _, _, _, e1, g1 = jobs[j]
        assert abs(e[j] - e1) <= 1e-12 * max(1, abs(e1))
        assert rel(g[j], g1) <= 1e-12
    for j in (1, 2, 4):                      # untouched circuits keep their results
        assert e[j] == e0[j]


def test_batch_against_oracle_shift(fsa, orc, circ):
    jobs = small_batch(circ, 13, 3, 10, 13, terms=5)
    b = fsa.Batch("fp64")
    for n, ops, coeffs, paulis in jobs:
        b.add(fsa.Circuit(ops[:40]), n, coeffs, paulis)
    e, g = b.run()
    for j, (n, ops, coeffs, paulis) in enumerate(jobs):
        ops = ops[:40]
        thetas = [a[-1] for nm, a in ops if nm == "pauliGadget"]

        def value(th):
            psi = np.zeros(1 << n, dtype=np.complex128)
            psi[0] = 1
            k = 0
            for nm, a in ops:
                if nm == "pauliGadget":
                    a = a[:-1] + (th[k],)
                    k += 1
                getattr(orc, "sv_" + nm)(psi, *a)
            return orc.sv_innerProduct(psi, orc.sv_applyPauliSum(psi, coeffs, paulis)).real

        assert abs(e[j] - value(thetas)) <= 1e-12
        shift = []
        for k in range(len(thetas)):
            tp, tm = list(thetas), list(thetas)
            tp[k] += np.pi / 4
            tm[k] -= np.pi / 4
            shift.append(value(tp) - value(tm))
        assert rel(g[j], shift) <= 1e-11


def test_batch_fp32_error_stated(fsa, circ):
    jobs = small_batch(circ, 14, 16, 10, 16)
    b32, b64 = fsa.Batch("fp32"), fsa.Batch("fp64")
    for n, ops, coeffs, paulis in jobs:
        c = fsa.Circuit(ops)
        b32.add(c, n, coeffs, paulis)
        b64.add(c, n, coeffs, paulis)
    e32, g32 = b32.run()
    e64, g64 = b64.run()
    ge_ = rel(np.concatenate(g32), np.concatenate(g64))
    ee = rel(e32, e64)
    print(f"batch FP32 vs FP64: <H> rel err {ee:.2e}, gradient rel err {ge_:.2e}")
    assert ge_ <= 1e-6 and ee <= 1e-6


def test_batch_autotune_keeps_results(fsa, circ):
    jobs = small_batch(circ, 15, 12, 10, 14)
    b = fsa.Batch("fp64")
    for n, ops, coeffs, paulis in jobs:
        b.add(fsa.Circuit(ops), n, coeffs, paulis)
    e0, g0 = b.run()
    ms = b.autotune(sample=6)
    assert ms > 0
    tb, m, low = b.config
    assert tb in (256, 512) and 10 <= m <= 12 and low in (4, 5, 6)
    e, g = b.run()
    assert rel(e, e0) <= 1e-12
    assert rel(np.concatenate(g), np.concatenate(g0)) <= 1e-12


def test_batch_errors(fsa, circ):
    b = fsa.Batch("fp64")
    with pytest.raises(fsa.FsaError):
        b.run()                               # empty
    c = fsa.Circuit([("oneTargGate", (12, np.eye(2)))])
    with pytest.raises(fsa.FsaError):
        b.add(c, 10, [1.0], np.zeros((1, 10)))     # qubit 12 >= 10
    with pytest.raises(fsa.FsaError):
        b.set_config(128, 11, 5)
    with pytest.raises(fsa.FsaError):
        b.set_config(256, 14, 5)


def test_batch_with_18_to_20_qubit_jobs(fsa, circ):
    """The benchmark batch's largest jobs (18-20 qubits, streamed through the
    arena tile by tile) among small LDS-resident ones, FP64 and FP32, against
    the single-state adjoint of the same precision's reference (FP64)."""
    pool = circ.config5(seed=16, num=160, nmin=10, nmax=20, num_terms=20)
    big = [j for j in pool if j[0] >= 18]
    jobs = []
    for n_ in (18, 19, 20):
        jobs += [j for j in big if j[0] == n_][:1]
    jobs += [j for j in pool if j[0] <= 13][:5]
    assert sorted({j[0] for j in jobs})[-3:] == [18, 19, 20]
    singles = [single_adjoint(fsa, n, ops, coeffs, paulis) for n, ops, coeffs, paulis in jobs]
    for precision, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        b = fsa.Batch(precision)
        for n, ops, coeffs, paulis in jobs:
            b.add(fsa.Circuit(ops), n, coeffs, paulis)
        e, g = b.run()
        errs = [rel(g[j], singles[j][1]) for j in range(len(jobs))]
        eerr = max(abs(e[j] - singles[j][0]) / max(1.0, abs(singles[j][0])) for j in range(len(jobs)))
        agg = rel(np.concatenate(g), np.concatenate([s[1] for s in singles]))
        print(f"batch {precision} with 18-20 qubit jobs: gradient rel err per job max {max(errs):.2e}, "
              f"aggregate {agg:.2e}, <H> {eerr:.2e}")
        assert agg <= tol and eerr <= tol, (precision, errs)


def test_batch_fp32_against_oracle_shift(fsa, orc, circ):
    # FP32 amplitudes, FP64 sums: gradients against the FP64 oracle's shift rule
    jobs = small_batch(circ, 17, 3, 10, 12, terms=5)
    b = fsa.Batch("fp32")
    for n, ops, coeffs, paulis in jobs:
        b.add(fsa.Circuit(ops[:60]), n, coeffs, paulis)
    e, g = b.run()
    for j, (n, ops, coeffs, paulis) in enumerate(jobs):
        ops = ops[:60]
        thetas = [a[-1] for nm, a in ops if nm == "pauliGadget"]

        def value(th):
            psi = np.zeros(1 << n, dtype=np.complex128)
            psi[0] = 1
            k = 0
            for nm, a in ops:
                if nm == "pauliGadget":
                    a = a[:-1] + (th[k],)
                    k += 1
                getattr(orc, "sv_" + nm)(psi, *a)
            return orc.sv_innerProduct(psi, orc.sv_applyPauliSum(psi, coeffs, paulis)).real

        shift = []
        for k in range(len(thetas)):
            tp, tm = list(thetas), list(thetas)
            tp[k] += np.pi / 4
            tm[k] -= np.pi / 4
            shift.append(value(tp) - value(tm))
        err = rel(g[j], shift)
        print(f"FP32 batch job {j} (n={n}): gradient rel err vs oracle shift {err:.2e}")
        assert abs(e[j] - value(thetas)) <= 1e-6 * max(1.0, abs(value(thetas)))
        assert err <= 1e-6
