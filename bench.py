#!/usr/bin/env python3
"""Benchmark: seeded 30-qubit FP64 statevector circuit of 1000 gates on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30] [--no-cpu-baseline]

One step applies the whole 1000-gate circuit (fused passes) to a 16 GiB state
already resident in HBM.  Prints ONE JSON line on stdout; progress goes to
stderr.  Fields beyond the common contract:
  roofline      -- the kernel with the most GPU time in the timed region:
                   algorithmic bytes per launch / mean launch time (HIP events
                   on the state's stream), against 8 TB/s;
  per_gate_type -- each gate type run as its own streaming kernel (eager
                   mode), algorithmic GB/s (2 x 16 B x amplitudes it changes);
  cpu_baseline  -- the CPU oracle (C99/OpenMP restatement of the reference's
                   loops) on the first gates of the same circuit, host cores;
                   per_config: the same oracle on bounded samples of the
                   12-qubit circuit, the 14-qubit density matrix, the 12-qubit
                   parameter-shift gradient and the 4096-circuit batch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__ as ge  # noqa: E402

METRIC = "HBM GB/s & gates/s @30q SV; adjoint grad circuits/s FP32 vs FP64 @ stated err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_circuits_module():
    import importlib.util

    spec = importlib.util.spec_from_file_location("fsa_amd_circuits", ge.PKG_DIR / "circuits.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def per_gate_type(fsa, circ, n, ops, reps=6):
    """Eager (one kernel per gate) timing per gate type on the n-qubit state."""
    psi = fsa.StateVector(n, mode=fsa.FSA_EXEC_EAGER)
    psi.init_gaussian(11)
    psi.sync()
    by_type = {}
    for name, args in ops:
        by_type.setdefault(name, [])
        if len(by_type[name]) < reps:
            by_type[name].append(args)
    out = {}
    for name, arglist in sorted(by_type.items()):
        fsa.apply_ops(psi, [(name, arglist[0])])     # warm the kernel
        psi.sync()
        fsa.profile_enable(True)
        for args in arglist:
            fsa.apply_ops(psi, [(name, args)])
        psi.sync()
        prof = fsa.profile_read()
        fsa.profile_enable(False)
        ms = sum(v[1] for v in prof.values())
        nbytes = sum(circ.algorithmic_bytes(name, a, n) for a in arglist)
        out[name] = {"gates": len(arglist), "ms_per_gate": round(ms / len(arglist), 4),
                     "GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    psi.close()
    return out


def cpu_baseline(circ, n, ops, seed, budget_s=15.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    orc = oracle.oracle()
    t0 = time.perf_counter()
    psi = orc.init_gaussian(n, seed)
    log(f"cpu baseline: {n}-qubit state initialised in {time.perf_counter() - t0:.1f}s")
    done, nbytes = 0, 0
    t0 = time.perf_counter()
    for name, args in ops:
        getattr(orc, "sv_" + name)(psi, *args)
        done += 1
        nbytes += circ.algorithmic_bytes(name, args, n)
        if time.perf_counter() - t0 > budget_s and done >= 3:
            break
    dt = time.perf_counter() - t0
    del psi
    return {"value": round(done / dt, 3), "unit": "gates/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"first {done} gates of the same {n}-qubit circuit on the C99/OpenMP oracle "
                      f"({dt:.1f} s, {nbytes / dt / 1e9:.1f} GB/s algorithmic)"}


def _orc():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    return oracle.oracle()


def _sv_value(orc, psi0, ops, coeffs, paulis, thetas=None):
    """<psi|H|psi> after ops on a copy of psi0 (gadget angles from thetas)."""
    psi = psi0.copy()
    k = 0
    for name, args in ops:
        if thetas is not None and name in ("pauliGadget", "phaseGadget"):
            args = args[:-1] + (thetas[k],)
            k += 1
        getattr(orc, "sv_" + name)(psi, *args)
    return orc.sv_innerProduct(psi, orc.sv_applyPauliSum(psi, coeffs, paulis)).real


def cpu_baselines_per_config(circ, jobs5=None, budget_s=6.0):
    """The oracle on the other workloads, each on a bounded sample (seconds of
    host time), measured or extrapolated as each entry's "sample" says."""
    orc = _orc()
    cores = orc.num_threads()
    out = {}
    # 12-qubit statevector, the whole 200-gate circuit
    ops1 = circ.config1()
    psi = orc.init_gaussian(12, 20261203)
    t0 = time.perf_counter()
    reps = 0
    while True:
        for name, args in ops1:
            getattr(orc, "sv_" + name)(psi, *args)
        reps += 1
        if time.perf_counter() - t0 > 1.0:
            break
    dt = (time.perf_counter() - t0) / reps
    out["sv12_200_gates"] = {"value": round(len(ops1) / dt, 1), "unit": "gates/s", "cores": cores, "kind": "port",
                             "sample": f"the whole 200-gate 12-qubit circuit, {reps} runs"}
    # 14-qubit density matrix: ops until the budget, expecPauliString over 1 and
    # 4 strings (64 extrapolated linearly), partialTrace over 4 qubits
    ops3, coeffs3, paulis3, tt3 = circ.config3()
    N = 14
    rho = orc.init_gaussian(2 * N, 20261205)
    t0 = time.perf_counter()
    done = 0
    for name, args in ops3:
        getattr(orc, "dm_" + name)(rho, *args)
        done += 1
        if time.perf_counter() - t0 > budget_s and done >= 3:
            break
    t_op = (time.perf_counter() - t0) / done
    pa = np.asarray(paulis3, dtype=np.uint32).reshape(64, N)
    t_e = {}
    for T in (1, 4):
        t0 = time.perf_counter()
        orc.dm_expecPauliString(rho, coeffs3[:T], pa[:T].ravel())
        t_e[T] = time.perf_counter() - t0
    per_term = max(t_e[4] - t_e[1], 0.0) / 3
    t_exp64 = t_e[1] + 63 * per_term
    t0 = time.perf_counter()
    orc.dm_partialTrace(rho, tt3)
    t_pt = time.perf_counter() - t0
    del rho
    out["dm14_400_ops"] = {"value": round(1.0 / t_op, 3), "unit": "ops/s", "cores": cores, "kind": "port",
                           "expecPauliString_64_s": round(t_exp64, 3), "partialTrace_4_s": round(t_pt, 3),
                           "sample": f"first {done} of the 400 ops ({done * t_op:.1f} s); expec over 1 and 4 "
                                     f"strings ({t_e[1]:.2f} s, {t_e[4]:.2f} s), 64 extrapolated linearly"}
    # 12-qubit copy of the layered gadget circuit: parameter-shift gradient =
    # 2 x angles evaluations of <H>; a few evaluations timed
    ops4, coeffs4, paulis4 = circ.config4(n=12)
    psi0 = np.zeros(1 << 12, dtype=np.complex128)
    psi0[0] = 1
    P = len(ops4)
    t0 = time.perf_counter()
    evals = 0
    while evals < 3 or time.perf_counter() - t0 < 1.0:
        _sv_value(orc, psi0, ops4, coeffs4, paulis4)
        evals += 1
    t_eval = (time.perf_counter() - t0) / evals
    out["adjoint12_parameter_shift"] = {
        "value": round(1.0 / ((2 * P + 1) * t_eval), 4), "unit": "gradients/s", "cores": cores, "kind": "port",
        "sample": f"{evals} evaluations of <H> on the 12-qubit copy ({P} angles, {len(coeffs4)} strings); "
                  f"a parameter-shift gradient is 2 x {P} + 1 of them"}
    # batch: forward <H> of a seeded sample of jobs; each job's <H> and
    # gradient by parameter shift costs (2 x gadgets + 1) forward evaluations
    if jobs5 is not None:
        rng = np.random.default_rng(5)
        order = rng.permutation(len(jobs5))
        est, t_meas, used = [], 0.0, 0
        for j in order:
            n, ops, coeffs, paulis = jobs5[j]
            psi0 = np.zeros(1 << n, dtype=np.complex128)
            psi0[0] = 1
            t0 = time.perf_counter()
            _sv_value(orc, psi0, ops, coeffs, paulis)
            t = time.perf_counter() - t0
            t_meas += t
            used += 1
            gad = sum(1 for nm, _ in ops if nm == "pauliGadget")
            est.append((2 * gad + 1) * t)
            if t_meas > budget_s and used >= 16:
                break
        total = float(np.mean(est)) * len(jobs5)
        out["batch4096_parameter_shift"] = {
            "value": round(len(jobs5) / total, 3), "unit": "circuits/s", "cores": cores, "kind": "port",
            "sample": f"forward <H> of {used} random jobs ({t_meas:.1f} s); each job's gradient costs "
                      f"2 x gadgets + 1 forward evaluations, mean scaled to {len(jobs5)} jobs"}
    return out


def adjoint_bench(fsa, circ, n=24, reps=5):
    """Layered ZZ + X gadget circuit from |0..0> (20 layers, 940 angles at 24
    qubits), H = 100 Pauli strings: <H> and the adjoint gradient of every angle,
    FP64 and FP32-bulk amplitudes (FP64 sums).  circuits/s counts whole
    forward + reverse sweeps; errors are FP32 against FP64 on the same GPU."""
    ops, coeffs, paulis = circ.config4(n=n)
    c = fsa.Circuit(ops)
    out = {"workload": f"{n}-qubit |0..0>, 20 layers of ZZ + X gadgets ({c.num_params} angles), "
                       f"H = {len(coeffs)} Pauli strings of weight 1-4"}
    res = {}
    for prec in ("fp64", "fp32"):
        psi0 = fsa.StateVector(n, precision=prec)
        psi0.init_zero()
        fsa.adjoint_gradient(c, psi0, coeffs, paulis)        # warm-up
        # timed without the per-launch HIP events of the library's profiler,
        # then one more call with them for the kernel breakdown
        t0 = time.perf_counter()
        for _ in range(reps):
            e, g = fsa.adjoint_gradient(c, psi0, coeffs, paulis)
        dt = (time.perf_counter() - t0) / reps
        fsa.profile_enable(True)
        fsa.adjoint_gradient(c, psi0, coeffs, paulis)
        prof = fsa.profile_read()
        fsa.profile_enable(False)
        psi0.close()
        res[prec] = (e, g)
        top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:6]
        out[prec] = {"circuits_per_s": round(1.0 / dt, 3), "ms_per_circuit": round(dt * 1e3, 2),
                     "kernel_ms_per_circuit": {k: round(v[1], 2) for k, v in top},
                     "launches_per_circuit": {k: round(v[0], 1) for k, v in top}}
        log(f"adjoint {prec}: {dt * 1e3:.1f} ms per circuit")
    e64, g64 = res["fp64"]
    e32, g32 = res["fp32"]
    out["fp32_speedup"] = round(out["fp32"]["circuits_per_s"] / out["fp64"]["circuits_per_s"], 3)
    out["fp32_grad_rel_err_vs_fp64"] = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    out["fp32_expec_rel_err_vs_fp64"] = float(abs(e32 - e64) / abs(e64))
    out["expec_fp64"] = e64
    return out


def dm_bench(fsa, circ, N=14, seed=20261205):
    """14-qubit density matrix (2^28 complex128 = 4 GiB), rank-8 random mixed
    state, 400 channel / unitary ops in fused passes, then expecPauliString
    over 64 strings and partialTrace over 4 qubits."""
    ops, coeffs, paulis, tt = circ.config3(seed=seed, N=N)
    rho_h = circ.random_mixed_state(N, seed)
    rho = fsa.DensityMatrix(N)
    rho.upload(rho_h)
    del rho_h
    fsa.apply_ops(rho, ops, dm=True)                  # warm-up (kernels, schedule)
    rho.sync()
    fsa.profile_enable(True)
    t0 = time.perf_counter()
    fsa.apply_ops(rho, ops, dm=True)
    rho.sync()
    t_ops = time.perf_counter() - t0
    prof = fsa.profile_read()
    fsa.profile_enable(False)
    fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)     # warm-up
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        ex = fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)
    t_exp = (time.perf_counter() - t0) / reps
    fsa.distributed_densitymatrix_partialTrace(rho, tt).close()
    t0 = time.perf_counter()
    for _ in range(reps):
        red = fsa.distributed_densitymatrix_partialTrace(rho, tt)
        red.sync()
        red.close()
    t_pt = (time.perf_counter() - t0) / reps
    rho.close()
    nbytes = sum(circ.algorithmic_bytes(nm, ar, 2 * N, dm=True) for nm, ar in ops)
    tiles = {k: v for k, v in prof.items() if k.startswith("tile_pass")}
    passes = sum(v[0] for v in tiles.values())
    tile_ms = sum(v[1] for v in tiles.values())
    pass_bytes = 2 * 16 * (1 << (2 * N))
    return {
        "workload": f"{N}-qubit density matrix (2^{2 * N} complex128), {len(ops)} ops "
                    "(manyTargGate k=1-3, 1/2-qubit dephasing and depolarising, damping), "
                    f"expecPauliString over {len(coeffs)} strings, partialTrace over {len(tt)} qubits",
        "ops_per_s": round(len(ops) / t_ops, 1),
        "ms_for_ops": round(t_ops * 1e3, 2),
        "fused_passes": passes,
        "tile_pass_GB_per_s": round(passes * pass_bytes / (tile_ms * 1e-3) / 1e9, 1) if passes else None,
        "effective_GB_per_s": round(nbytes / t_ops / 1e9, 1),
        "kernel_ms": {k: round(v[1], 2) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "expec_ms": round(t_exp * 1e3, 3),
        "expec_us_per_term": round(t_exp * 1e6 / len(coeffs), 2),
        "expec_value": [ex.real, ex.imag],
        "partial_trace_ms": round(t_pt * 1e3, 3),
    }


def batch_bench(fsa, circ, jobs, reps=3, sample=512):
    """4096 independent circuits (n log-uniform in [10, 20], 50-300 gates,
    each with its own 20-string observable): every <H> and every gadget
    gradient from one persistent-kernel graph launch, FP64 and FP32, with the
    configuration picked by the batch autotuner on a spread sample."""
    t0 = time.perf_counter()
    circuits = [fsa.Circuit(ops) for _, ops, _, _ in jobs]
    log(f"batch: {len(jobs)} circuits generated in {time.perf_counter() - t0:.1f} s")
    ns = np.array([j[0] for j in jobs])
    gates = sum(len(j[1]) for j in jobs)
    out = {"workload": f"{len(jobs)} statevectors from |0..0>, n log-uniform in [10, 20] (mean {ns.mean():.2f}), "
                       f"{gates} gates in total (oneTarg, 1-3 controls, manyTarg k=2-3, 1-4 qubit gadgets), "
                       f"20-string observable each; forward <H> + adjoint gradient of every gadget",
           "circuits": len(jobs), "gates": gates}
    res = {}
    for prec in ("fp64", "fp32"):
        b = fsa.Batch(prec)
        for (n, _, coeffs, paulis), c in zip(jobs, circuits):
            b.add(c, n, coeffs, paulis)
        t1 = time.perf_counter()
        tune_ms = b.autotune(sample=sample)
        t_tune = time.perf_counter() - t1
        b.run()                                   # build, capture, first replay
        fsa.profile_enable(True)
        t1 = time.perf_counter()
        for _ in range(reps):
            e, g = b.run()
        dt = (time.perf_counter() - t1) / reps
        prof = fsa.profile_read()
        fsa.profile_enable(False)
        kname = f"batch_{prec}"
        kms = prof[kname][1] / prof[kname][0] if kname in prof else None
        res[prec] = (e, np.concatenate(g))
        out[prec] = {"circuits_per_s": round(len(jobs) / dt, 1), "ms_per_batch": round(dt * 1e3, 2),
                     "kernel_ms": round(kms, 2) if kms else None,
                     "gates_per_s": round(gates / dt, 1),
                     "config": {"workgroup": b.config[0], "tile_bits": b.config[1], "low_bits": b.config[2]},
                     "autotune_s": round(t_tune, 1), "autotune_sample_ms": round(tune_ms, 2)}
        log(f"batch {prec}: {dt * 1e3:.1f} ms per batch, config {b.config}")
        b.close()
    e64, g64 = res["fp64"]
    e32, g32 = res["fp32"]
    out["fp32_speedup"] = round(out["fp32"]["circuits_per_s"] / out["fp64"]["circuits_per_s"], 3)
    out["fp32_grad_rel_err_vs_fp64"] = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    out["fp32_expec_max_abs_err_vs_fp64"] = float(np.abs(e32 - e64).max())
    return out


def micro_passes(fsa, n, reps=3):
    """Two single-pass workloads on an n-qubit FP64 state, `reps` launches each:
    a pass holding one 5-target dense gate (MFMA tile variant) and a pass of
    two phase gadgets (plain variant).  For hardware-counter runs."""
    rng = np.random.default_rng(1)
    z = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    u, _, vh = np.linalg.svd(z)
    hi = list(range(n - 7, n))
    psi = fsa.StateVector(n)
    psi.init_gaussian(3)
    dense = fsa.Circuit([("manyTargGate", (hi[:5], u @ vh)), ("phaseGadget", ([1, 2], 0.3))])
    plain = fsa.Circuit([("phaseGadget", ([1, 20], 0.3)), ("phaseGadget", ([3, 22], 0.4))])
    fsa.profile_enable(True)
    for c in (dense, plain):
        for _ in range(reps):
            c.apply(psi)
        psi.sync()
    prof = fsa.profile_read()
    fsa.profile_enable(False)
    print(json.dumps({k: {"launches": v[0], "mean_ms": round(v[1] / v[0], 4)} for k, v in prof.items()}), flush=True)


def dense_gates(fsa, n=28, reps=3):
    """Dense gates of 6-12 targets on an n-qubit state (n - 2 from 9 targets),
    eager, FP64 and FP32: the MFMA GEMM (kernels_gemm.hip) and, at the same
    sizes, the per-group kernel it replaces (FSA_GEMM_MINK=13).  ms per gate
    from the library's HIP events, flops 8 * 2^k * 2^n."""
    rng = np.random.default_rng(1)
    out = []
    for prec in ("fp64", "fp32"):
        for k in range(6, 13):
            nq = n if k <= 8 else n - 2
            d = 1 << k
            g = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
            t = [int(x) for x in rng.permutation(nq)[:k]]
            row = {"precision": prec, "k": k, "qubits": nq}
            for form, mink in (("gemm", "6"), ("per_group", "13")):
                if form == "per_group" and k >= 10:
                    continue            # seconds per gate at these sizes
                os.environ["FSA_GEMM_MINK"] = mink
                psi = fsa.StateVector(nq, precision=prec, mode=fsa.FSA_EXEC_EAGER)
                psi.init_gaussian(3)
                fsa.distributed_statevector_manyTargGate(psi, t, g)
                psi.sync()
                fsa.profile_enable(True)
                for _ in range(reps):
                    fsa.distributed_statevector_manyTargGate(psi, t, g)
                psi.sync()
                prof = fsa.profile_read()
                fsa.profile_enable(False)
                psi.close()
                ms = prof["matk"][1] / prof["matk"][0]
                row[form + "_ms"] = round(ms, 3)
                row[form + "_TFLOP_per_s"] = round(8 * (1 << (nq + k)) / (ms * 1e-3) / 1e12, 2)
            os.environ.pop("FSA_GEMM_MINK", None)
            out.append(row)
            print(json.dumps(row), flush=True)
    return out


def copy_calibration(n, reps=5):
    """A plain device-to-device copy of one n-qubit FP64 state (torch copy_),
    timed with events: the practical streaming ceiling next to the 8 TB/s peak,
    and the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    nbytes = 16 << n
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    src.normal_()
    dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_copy": round(ms, 4),
                      "GB_per_s": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--gates", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-type", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-dm", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only time a plain copy of one state-sized buffer and exit")
    ap.add_argument("--micro-passes", action="store_true",
                    help="only run one dense-gate pass and one diagonal pass (3 launches each) and exit")
    ap.add_argument("--dense-gates", action="store_true",
                    help="only time dense gates of 6-12 targets (GEMM and per-group kernels) and exit")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU exchange "
              "layer is not part of this build", file=sys.stderr)
        sys.exit(2)

    if a.calibrate_copy:
        copy_calibration(a.qubits)
        return

    fsa = ge.load_package()
    if a.micro_passes:
        micro_passes(fsa, a.qubits)
        return
    if a.dense_gates:
        dense_gates(fsa, min(a.qubits, 28))
        return
    circ = load_circuits_module()
    n = a.qubits
    seed = 20261204
    ops = circ.config2(seed=seed, n=n, num_gates=a.gates)
    total_bytes = sum(circ.algorithmic_bytes(nm, ar, n) for nm, ar in ops)

    log(f"building {len(ops)}-gate circuit for {n} qubits")
    c = fsa.Circuit(ops)
    passes = c.num_passes(n, "fp64")
    log(f"fused schedule: {passes} HBM passes for {len(ops)} gates")

    psi = fsa.StateVector(n)
    psi.init_gaussian(seed)
    fsa.comm_synch()
    for w in range(a.warmup):
        c.apply(psi)
        psi.sync()
        log(f"warmup step {w + 1}/{a.warmup} done")

    fsa.comm_synch()
    fsa.profile_enable(True)
    t0 = time.perf_counter()
    for k in range(a.steps):
        c.apply(psi)
        psi.sync()
        log(f"step {k + 1}/{a.steps} done")
    fsa.comm_synch()
    dt = time.perf_counter() - t0
    prof = fsa.profile_read()
    fsa.profile_enable(False)
    norm = psi.norm2()
    psi.close()

    gates_per_s = len(ops) * a.steps / dt
    ms_per_step = dt / a.steps * 1e3
    dominant = max(prof.items(), key=lambda kv: kv[1][1])
    dom_name, (dom_count, dom_ms) = dominant
    mean_ms = dom_ms / dom_count
    bytes_per_launch = 2 * 16 * (1 << n)        # every fused pass / full-state op reads and writes all amps
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(f"{dom_name}_{n}q_bytes_per_launch")
        except Exception:
            traffic = None

    # whole-step floor: every pass moves the state once at the HBM peak, and the
    # dense gates of 3-5 targets need their FP64 MFMA time at the matrix peak
    mfma_flops = sum(8 * (1 << len(a[0])) * (1 << n) for nm, a in ops if nm == "manyTargGate" and len(a[0]) >= 3)
    hbm_floor_ms = passes * bytes_per_launch / (HBM_PEAK_GBS * 1e9) * 1e3
    mfma_floor_ms = mfma_flops / (FP64_PEAK_TFLOPS * 1e12) * 1e3
    result = {
        "metric": METRIC,
        "value": round(gates_per_s, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{n}-qubit FP64 statevector, seeded {len(ops)}-gate circuit (oneTarg, 2-control, "
                        f"swap, 3-qubit pauliGadget, manyTarg k=2..5, 4-qubit phaseGadget, 6-qubit pauliTensor)",
            "qubits": n,
            "gates": len(ops),
            "fused_passes_per_step": passes,
            "algorithmic_GB_per_step": round(total_bytes / 1e9, 1),
            "effective_GB_per_s": round(total_bytes * a.steps / dt / 1e9, 1),
            "final_norm2": norm,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": dom_name,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "launches": dom_count,
            "mean_launch_ms": round(mean_ms, 4),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "step_hbm_floor_ms": round(hbm_floor_ms, 1),
            "step_mfma_floor_ms": round(mfma_floor_ms, 1),
            "step_frac_of_floor": round(max(hbm_floor_ms, mfma_floor_ms) / ms_per_step, 4),
        },
        "kernel_time_ms": {k: round(v[1], 2) for k, v in prof.items()},
    }
    if not a.no_adjoint:
        log("adjoint gradient, FP64 and FP32-bulk")
        result["adjoint"] = adjoint_bench(fsa, circ)
    if not a.no_dm:
        log("density matrix: 400 ops, expecPauliString, partialTrace")
        result["density_matrix"] = dm_bench(fsa, circ)
    jobs5 = None
    if not a.no_batch or not a.no_cpu_baseline:
        t0 = time.perf_counter()
        jobs5 = circ.config5()
        log(f"batch: {len(jobs5)} circuits generated in {time.perf_counter() - t0:.1f} s")
    if not a.no_batch:
        log("batch: 4096 circuits with gradients")
        result["batch"] = batch_bench(fsa, circ, jobs5)
    if not a.no_per_type:
        log("per-gate-type eager timing")
        result["per_gate_type"] = per_gate_type(fsa, circ, n, ops)
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle) on a bounded sample")
        result["cpu_baseline"] = cpu_baseline(circ, n, ops, seed)
        log("cpu baselines of the other workloads")
        result["cpu_baseline"]["per_config"] = cpu_baselines_per_config(circ, jobs5)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
