"""libfsa's kernel source run on the CPU under AddressSanitizer.

tests/emu/ holds a CPU stand-in for the HIP runtime and device built-ins (one
OS thread per lane, real barriers, wave exchanges, the 16x16x4 MFMA with the
operand layout measured on gfx950).  The real planner (engine.cpp), the fused
tile kernel (kernels_tile.hip), the streaming kernels that run passes of one
op (kernels_stream.hip) and the batch kernel (kernels_batch.hip) are
compiled against it with clang++ -fsanitize=address, so every LDS and global
access of those kernels is bounds-checked on the host, and their results are
compared with the oracle:

  * fused passes of 3-5 target dense gates on 2^12- and 2^13-amplitude tiles
    (MFMA form with the matrices staged in LDS, and VALU form; the VALU split
    mode for 2^13 FP64 tiles was wrong before it was restructured into rounds
    of whole groups);
  * a pass whose last item is a phase gadget applied in registers as each
    tile is written back;
  * a mixed random circuit on 2^13-amplitude tiles;
  * batched <H> and adjoint gradients, LDS-resident and streamed jobs, FP64 and
    FP32, against an adjoint computed with the oracle;
  * the single-state adjoint sweep (kernels_reduce.hip): runs of one-qubit X/Y
    gadgets (group and lane forms, a diagonal tail), runs of Z-only gadgets and
    single steps, FP64 and FP32, against the same oracle adjoint;
  * every op kind as its own streaming kernel, on grids that take the
    XCD-contiguous block order and on grids that keep the round-robin one.

Skipped when the ROCm clang++ is not installed.
"""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"
CSRC = ROOT / "distributed-full-state-algorithms_amd" / "csrc"
BUILD = EMU / "_build"
CXX = os.environ.get("FSA_EMU_CXX", "/opt/rocm/llvm/bin/clang++")
FLAGS = ["-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-D_GNU_SOURCE",
         f"-I{EMU}", f"-I{ROOT / 'include'}", f"-I{CSRC}"]

CODES = {"oneTargGate": 0, "manyCtrlOneTargGate": 1, "swapGate": 2, "manyTargGate": 3, "pauliTensor": 4,
         "pauliGadget": 5, "phaseGadget": 6}


def _build():
    BUILD.mkdir(exist_ok=True)
    exe = BUILD / "emu_drive"
    sources = [CSRC / f for f in ("kernels_tile.hip", "kernels_stream.hip", "kernels_batch.hip", "kernels_reduce.hip",
                                  "kernels_rtile.hip", "kernels_gemm.hip", "rtile_plan.cpp", "engine.cpp",
                                  "circuit.cpp", "batch.cpp", "engine.hpp", "tile_ops.hpp", "common.hpp", "launch.hpp")]
    sources += [EMU / "emu_rt.cpp", EMU / "emu_drive.cpp", EMU / "hip" / "hip_runtime.h"]
    if exe.exists() and all(s.stat().st_mtime < exe.stat().st_mtime for s in sources):
        return exe
    # dynamic LDS becomes a static array (workgroups run one at a time)
    for name in ("kernels_tile", "kernels_stream", "kernels_rtile"):
        src = (CSRC / f"{name}.hip").read_text().replace(
            "extern __shared__ unsigned char smem_raw[];",
            "static unsigned char smem_raw[262144] __attribute__((aligned(16)));")
        (BUILD / f"{name}_emu.cpp").write_text(src)
    units = [(BUILD / "kernels_tile_emu.cpp", []), (BUILD / "kernels_stream_emu.cpp", []),
             (BUILD / "kernels_rtile_emu.cpp", []), (CSRC / "rtile_plan.cpp", []),
             (CSRC / "kernels_batch.hip", ["-x", "c++"]), (CSRC / "kernels_reduce.hip", ["-x", "c++"]),
             (CSRC / "kernels_gemm.hip", ["-x", "c++"]),
             (CSRC / "engine.cpp", []), (CSRC / "circuit.cpp", []), (CSRC / "batch.cpp", []),
             (EMU / "emu_rt.cpp", []), (EMU / "emu_drive.cpp", [])]
    procs, objs = [], []
    for src, extra in units:
        obj = BUILD / (src.stem + ".o")
        objs.append(obj)
        procs.append(subprocess.Popen([CXX, *FLAGS, *extra, "-c", str(src), "-o", str(obj)],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(out)
    # host launchers of kernels not emulated here are never called by the driven paths
    subprocess.run([CXX, "-fsanitize=address", *map(str, objs), "-o", str(exe), "-lpthread",
                    "-Wl,--unresolved-symbols=ignore-all"], check=True)
    return exe


@pytest.fixture(scope="module")
def drive():
    if not (Path(CXX).exists() or shutil.which(CXX)):
        pytest.skip("ROCm clang++ not installed")
    import fcntl

    BUILD.mkdir(exist_ok=True)
    # one build at a time when several test processes start together
    with open(BUILD / ".lock", "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        try:
            return _build()
        finally:
            fcntl.flock(lock, fcntl.LOCK_UN)


def _run(exe, args, env=None):
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    e.update(env or {})
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, env=e, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(kv.split("=") for kv in r.stdout.split() if "=" in kv)
    return {k: float(v) for k, v in vals.items()}


def _write_op(f, name, args):
    code = CODES[name]
    th, mat, p = 0.0, None, []
    if code == 0:
        q, mat = [args[0]], args[1]
    elif code == 1:
        q, mat = list(args[0]) + [args[1]], args[2]
    elif code == 2:
        q = [args[0], args[1]]
    elif code == 3:
        q, mat = list(args[0]), args[1]
    elif code == 4:
        q, p = list(args[0]), list(args[1])
    elif code == 5:
        q, p, th = list(args[0]), list(args[1]), args[2]
    else:
        q, th = list(args[0]), args[1]
    np.array([code, len(q)] + q + p, np.uint32).tofile(f)
    np.array([th], np.float64).tofile(f)
    if mat is not None:
        np.ascontiguousarray(np.asarray(mat, np.complex128)).tofile(f)


def _unitary(rng, d):
    u, _, vh = np.linalg.svd(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return u @ vh


def _random_op(rng, n, dense_only=False):
    q = [int(x) for x in rng.permutation(n)]
    kind = "many" if dense_only else ["one", "ctrl", "swap", "many", "tensor", "gadget", "phase"][int(rng.integers(7))]
    if kind == "many":
        k = int(rng.integers(3, 6)) if dense_only else int(rng.integers(2, 6))
        return "manyTargGate", (q[:k], _unitary(rng, 1 << k))
    if kind == "one":
        return "oneTargGate", (q[0], _unitary(rng, 2))
    if kind == "ctrl":
        return "manyCtrlOneTargGate", (q[1:3], q[0], _unitary(rng, 2))
    if kind == "swap":
        return "swapGate", (q[0], q[1])
    k = int(rng.integers(1, 5))
    p = [int(x) for x in rng.integers(1, 4, k)]
    if kind == "tensor":
        return "pauliTensor", (q[:k], p)
    if kind == "gadget":
        return "pauliGadget", (q[:k], p, float(rng.uniform(0, 6.3)))
    return "phaseGadget", (q[:k], float(rng.uniform(0, 6.3)))


def _circuit_file(path, orc, n, ops, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    ref = v.copy()
    for name, args in ops:
        getattr(orc, "sv_" + name)(ref, *args)
    with open(path, "wb") as f:
        np.array([n, len(ops)], np.uint32).tofile(f)
        for name, args in ops:
            _write_op(f, name, args)
        v.astype(np.complex128).tofile(f)
        ref.astype(np.complex128).tofile(f)


@pytest.mark.parametrize("bits", ["12", "13"])
@pytest.mark.parametrize("mfma", ["1", "0"])
def test_dense_passes_on_cpu(drive, orc, tmp_path, bits, mfma):
    # passes full of dense gates hold more than the 16 KiB of MFMA matrices a
    # workgroup keeps in LDS, so they take the VALU form
    rng = np.random.default_rng(31)
    n = 14
    path = tmp_path / "dense.bin"
    _circuit_file(path, orc, n, [_random_op(rng, n, dense_only=True) for _ in range(24)], rng)
    out = _run(drive, ["circuit", path], {"FSA_TILE_BITS": bits, "FSA_MFMA": mfma})
    assert out["rel_err"] <= 1e-12, out


@pytest.mark.parametrize("bits", ["12", "13"])
@pytest.mark.parametrize("ks", [(3,), (4,), (5,), (4, 4, 3), (5, 4), (5, 3, 5)])
def test_mfma_pass_with_matrices_in_lds_on_cpu(drive, orc, tmp_path, bits, ks):
    # one pass: dense gates among one-qubit gates, on MFMA with the matrices in
    # LDS; up to 16 KiB of FP64 matrices (32 KiB on 13-bit tiles) all of them
    # are staged once for the pass, beyond that each one just before its item
    rng = np.random.default_rng(33 + sum(ks))
    n = 14
    hi = [int(x) for x in rng.permutation(np.arange(4, n))]
    ops = [("oneTargGate", (int(rng.integers(n)), _unitary(rng, 2))) for _ in range(6)]
    for i, k in enumerate(ks):
        q = [int(x) for x in rng.permutation(hi[:8])[:k - 1]] + [i]     # within 9 tile qubits
        ops.append(("manyTargGate", (q, _unitary(rng, 1 << k))))
        ops.append(("oneTargGate", (q[0], _unitary(rng, 2))))
    path = tmp_path / "mfma.bin"
    _circuit_file(path, orc, n, ops, rng)
    # the LDS tile kernel (a remap-free dense pass may otherwise run register-resident)
    out = _run(drive, ["circuit", path], {"FSA_TILE_BITS": bits, "FSA_RTILE": "0"})
    assert out["rel_err"] <= 1e-12, out
    mat_bytes = sum(16 << (2 * k) for k in ks)
    per_pass = mat_bytes <= (32768 if bits == "13" else 16384)
    name = "tile_pass_mfma" if per_pass else "tile_pass_mfma_staged"
    assert out.get(name, 0) == 1 and out["passes"] == 1, out


@pytest.mark.parametrize("drain", ["1", "0"])
def test_phase_gadget_at_pass_end_on_cpu(drive, orc, tmp_path, drain):
    # one pass whose phase gadgets commute with the ops after them: one of them
    # is applied in registers as each tile is written back (FSA_DRAIN_DIAG=1)
    rng = np.random.default_rng(35)
    n = 15
    ops = [("oneTargGate", (q, _unitary(rng, 2))) for q in (3, 9, 12, 14)]
    ops += [("phaseGadget", ([0, 5, 13], 0.7)), ("oneTargGate", (7, _unitary(rng, 2))),
            ("phaseGadget", ([2, 9, 14], -1.1)), ("oneTargGate", (11, _unitary(rng, 2)))]
    path = tmp_path / "drain.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_DRAIN_DIAG": drain})
    assert out["rel_err"] <= 1e-12 and out["passes"] == 1, out


@pytest.mark.parametrize("drain", ["1", "0"])
def test_z_pauli_pair_at_pass_end_on_cpu(drive, orc, tmp_path, drain):
    # Z-only Pauli gadgets and tensors are one-mask diagonals too: two of the
    # three diagonals that commute to the pass end share its register item
    # (factors multiplied), the third stays an LDS item; one with an X does not
    # commute past the gate on its qubit and stays in place
    rng = np.random.default_rng(36)
    n = 15
    ops = [("oneTargGate", (q, _unitary(rng, 2))) for q in (3, 9, 12, 14)]
    ops += [("pauliGadget", ([1, 6, 13], [3, 3, 3], 0.4)), ("oneTargGate", (7, _unitary(rng, 2))),
            ("pauliTensor", ([0, 10], [3, 3])), ("pauliGadget", ([4, 7], [1, 3], 0.9)),
            ("oneTargGate", (4, _unitary(rng, 2))), ("phaseGadget", ([2, 9, 14], -1.1)),
            ("oneTargGate", (11, _unitary(rng, 2)))]
    path = tmp_path / "drainz.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_DRAIN_DIAG": drain})
    assert out["rel_err"] <= 1e-12 and out["passes"] == 1, out


def test_mixed_circuit_on_13_bit_tiles_on_cpu(drive, orc, tmp_path):
    rng = np.random.default_rng(32)
    n = 15
    ops = [op for op in (_random_op(rng, n) for _ in range(80)) if op[0] != "manyCtrlOneTargGate"]
    path = tmp_path / "mixed.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_TILE_BITS": "13"})
    assert out["rel_err"] <= 1e-12, out


def _oracle_adjoint(orc, n, ops, coeffs, paulis):
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    for name, a in ops:
        getattr(orc, "sv_" + name)(psi, *a)
    lam = orc.sv_applyPauliSum(psi, coeffs, paulis)
    expec = orc.sv_innerProduct(psi, lam).real
    grads, phi = [], psi
    for name, a in reversed(ops):
        if name == "pauliGadget":
            t, p, th = a
            pphi = phi.copy()
            orc.sv_pauliTensor(pphi, t, p)
            grads.append(-2 * np.vdot(lam, pphi).imag)
            inv = (t, p, -th)
        elif name == "oneTargGate":
            inv = (a[0], a[1].conj().T)
        elif name == "manyCtrlOneTargGate":
            inv = (a[0], a[1], a[2].conj().T)
        else:
            inv = (a[0], a[1].conj().T)
        getattr(orc, "sv_" + name)(phi, *inv)
        getattr(orc, "sv_" + name)(lam, *inv)
    return expec, grads[::-1]


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_adjoint_runs_on_cpu(drive, orc, tmp_path, prec, tol):
    # fsa_adjoint_gradient on 9 qubits: per layer a run of 10 Z-only gadgets, then
    # X/Y gadgets on qubits 0..8.  In the reverse sweep: qubits 8..6 take the
    # group kernel, 5..3 the lane kernel, 2..0 the lane kernel with the Z run's
    # first 8 gadgets as its tail, the rest of the Z run the diagonal kernel;
    # an X (x) Z gadget and the one-qubit gates take single steps
    rng = np.random.default_rng(52)
    n = 9
    ang = lambda: float(rng.uniform(0, 6.3))
    ops = []
    for layer in range(2):
        for i in range(n - 1):
            ops.append(("pauliGadget", ([i, i + 1], [3, 3], ang())))
        ops.append(("pauliGadget", ([0, 4, 8], [3, 3, 3], ang())))
        ops.append(("pauliGadget", ([1, 5], [3, 3], ang())))
        for q in range(n):
            ops.append(("pauliGadget", ([q], [1 if q % 3 else 2], ang())))
        ops.append(("pauliGadget", ([2, 7], [1, 3], ang())))
        ops.append(("oneTargGate", (4, _unitary(rng, 2))))
    T = 6
    paulis = np.zeros((T, n), np.uint32)
    for t in range(T):
        w = int(rng.integers(1, 4))
        paulis[t, rng.choice(n, w, replace=False)] = rng.integers(1, 4, w)
    coeffs = rng.normal(size=T)
    expec, grads = _oracle_adjoint(orc, n, ops, coeffs, paulis)
    path = tmp_path / "adjoint.bin"
    with open(path, "wb") as f:
        np.array([n, len(ops), T], np.uint32).tofile(f)
        for name, args in ops:
            _write_op(f, name, args)
        np.asarray(coeffs, np.float64).tofile(f)
        np.asarray(paulis, np.uint32).tofile(f)
        np.array([expec] + grads, np.float64).tofile(f)
    out = _run(drive, ["adjoint", path, prec])
    assert out["expec_abs_err"] <= tol and out["grad_rel_err"] <= tol, out
    assert out["adjoint_qrun"] == 6 and out["adjoint_diag"] == 2 and out["adjoint_step"] == 2, out


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_batch_kernel_on_cpu(drive, orc, tmp_path, prec, tol):
    import importlib.util

    spec = importlib.util.spec_from_file_location("fsa_amd_circuits",
                                                  ROOT / "distributed-full-state-algorithms_amd" / "circuits.py")
    circ = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(circ)
    jobs = circ.config5(seed=41, num=5, nmin=10, nmax=14, num_terms=5)
    path = tmp_path / "batch.bin"
    with open(path, "wb") as f:
        np.array([len(jobs)], np.uint32).tofile(f)
        for n, ops, coeffs, paulis in jobs:
            ops = ops[:40]
            expec, grads = _oracle_adjoint(orc, n, ops, coeffs, paulis)
            np.array([n, len(ops), len(coeffs)], np.uint32).tofile(f)
            for name, args in ops:
                _write_op(f, name, args)
            np.asarray(coeffs, np.float64).tofile(f)
            np.asarray(paulis, np.uint32).tofile(f)
            np.array([expec] + grads, np.float64).tofile(f)
    # tile bits 11: the 12..14-qubit jobs stream through LDS, the others stay in it
    out = _run(drive, ["batch", path, 512, 11, 4, prec])
    assert out["expec_abs_err"] <= tol and out["grad_rel_err"] <= tol, out


def test_standalone_dense_gate_between_passes_on_cpu(drive, orc, tmp_path):
    # an 8-target gate fits no tile, so it runs as a standalone kernel between
    # two tile passes and pushes its 1 MiB matrix into the parameter ring that
    # still holds the second pass's items (pushed before any launch): the ring
    # must keep them alive, where it once freed them (use after free)
    rng = np.random.default_rng(36)
    n = 12
    ops = [("oneTargGate", (3, _unitary(rng, 2))), ("oneTargGate", (9, _unitary(rng, 2))),
           ("manyTargGate", ([5, 0, 7, 2, 11, 3, 9, 6], _unitary(rng, 256))),
           ("oneTargGate", (0, _unitary(rng, 2))), ("oneTargGate", (7, _unitary(rng, 2)))]
    path = tmp_path / "ring.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path])
    assert out["rel_err"] <= 1e-12 and out["passes"] == 3 and out["tiled"] == 2, out


@pytest.mark.parametrize("k,n,eager,fp32,g3", [(6, 12, 1, 0, 1), (6, 13, 0, 1, 1), (7, 12, 1, 1, 1),
                                                (8, 12, 0, 0, 1), (9, 14, 0, 1, 1), (7, 13, 1, 0, 0)])
def test_dense_gate_gemm_on_cpu(drive, orc, tmp_path, k, n, eager, fp32, g3):
    # kernels_gemm.hip: in place for k <= 8, row panels through the scratch
    # area above (a 1 KiB cap: one column block per batch, so several
    # batches); targets in shuffled order so that low qubits are among them.
    # g3: three real products per complex one (FSA_GEMM_3M, the default) or four
    rng = np.random.default_rng(70 + k)
    q = [int(x) for x in rng.permutation(n)[:k]]
    ops = [("manyTargGate", (q, _unitary(rng, 1 << k))), ("oneTargGate", (q[0], _unitary(rng, 2)))]
    if k <= 8:
        ops.append(("manyTargGate", (sorted(q)[::-1], _unitary(rng, 1 << k))))
    path = tmp_path / "gemm.bin"
    _circuit_file(path, orc, n, ops, rng)
    env = {"FSA_GEMM_SCRATCH_KB": "1", "FSA_GEMM_3M": str(g3)}
    if eager:
        env["EMU_EAGER"] = "1"
    if fp32:
        env["EMU_FP32"] = "1"
    out = _run(drive, ["circuit", path], env)
    assert out["rel_err"] <= (2e-6 if fp32 else 1e-12), out


@pytest.mark.parametrize("n", [12, 15])
def test_eager_streaming_kernels_on_cpu(drive, orc, tmp_path, n):
    # every op as its own streaming kernel (kernels_stream.hip).  At 15 qubits
    # most grids are multiples of 8 blocks, which take the XCD-contiguous block
    # order; at 12 qubits (and for 2-control gates at 15) they have 1-4 blocks
    # and keep the round-robin one.  Both must cover every item exactly once.
    rng = np.random.default_rng(90 + n)
    ops = [_random_op(rng, n) for _ in range(40)]
    path = tmp_path / "eager.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"EMU_EAGER": "2"})
    assert out["rel_err"] <= 1e-12, out
    assert out.get("mat1", 0) > 0 and out.get("pauli", 0) + out.get("phase", 0) > 0, out


# ---- register-resident passes (kernels_rtile.hip, rtile_plan.cpp) ---------------

@pytest.mark.parametrize("rtile", ["1", ""])
@pytest.mark.parametrize("first", [[0, 1, 14, 2, 8], [2, 0, 1]])
def test_register_resident_dense_gate_on_low_qubits_first_on_cpu(drive, orc, tmp_path, rtile, first):
    # a pass whose first op is a dense gate on qubits 0, 1 and 2: those must sit
    # on lane slots for the load, but only two MFMA slots are lane slots, so
    # the pass loads in the identity layout and remaps first (the planner once
    # failed with "an op has no layout"); forced and automatic choice
    rng = np.random.default_rng(517)
    n = 16
    ops = [("manyTargGate", (first, _unitary(rng, 1 << len(first)))),
           ("manyTargGate", ([14, 8, 15, 6, 13], _unitary(rng, 32))),
           ("oneTargGate", (1, _unitary(rng, 2)))]
    path = tmp_path / "low.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_RTILE": rtile} if rtile else {})
    assert out["rel_err"] <= 1e-12, out
    if rtile:
        assert out.get("rtile_pass_mfma", 0) + out.get("rtile_pass_plain", 0) >= 1, out

@pytest.mark.parametrize("seed", [61, 62, 63, 64])
def test_register_resident_passes_on_cpu(drive, orc, tmp_path, seed):
    # random circuits on 12-bit FP64 tiles: one-qubit gates on lane and register
    # slots, controls anywhere, Pauli strings, phase gadgets and dense gates of
    # 2-5 targets on MFMA with operands from registers, with remaps between
    # layouts; swaps become relabelings whose fix-up passes run on the LDS kernel
    rng = np.random.default_rng(seed)
    n = 15
    ops = [_random_op(rng, n) for _ in range(70)]
    path = tmp_path / "rt.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_RTILE": "1"})
    assert out["rel_err"] <= 1e-12, out
    assert out.get("rtile_pass_mfma", 0) + out.get("rtile_pass_plain", 0) >= 2, out


def test_register_resident_remaps_on_cpu(drive, orc, tmp_path):
    # dense gates on qubits 0-2 (they leave the lane slots the HBM accesses need,
    # so the pass ends with a remap), Pauli strings over 9 qubits (X/Y on wave
    # slots), controls on wave slots, gates of every size in one pass
    rng = np.random.default_rng(65)
    n = 14
    ops = [("manyTargGate", ([0, 7, 2, 11, 5], _unitary(rng, 32))),
           ("oneTargGate", (9, _unitary(rng, 2))),
           ("pauliGadget", ([0, 1, 2, 3, 5, 7, 9, 11, 13], [1, 2, 1, 2, 1, 2, 1, 2, 3], 0.37)),
           ("manyCtrlOneTargGate", ([13, 12, 3], 6, _unitary(rng, 2))),
           ("manyTargGate", ([1, 12], _unitary(rng, 4))),
           ("manyTargGate", ([13, 0, 6], _unitary(rng, 8))),
           ("phaseGadget", ([1, 4, 8, 12], -0.8)),
           ("manyTargGate", ([3, 6, 9, 12], _unitary(rng, 16))),
           ("pauliTensor", ([2, 4, 6, 8, 10, 12], [1, 1, 2, 2, 3, 1])),
           ("oneTargGate", (1, _unitary(rng, 2)))]
    path = tmp_path / "remap.bin"
    _circuit_file(path, orc, n, ops, rng)
    out = _run(drive, ["circuit", path], {"FSA_RTILE": "1"})
    assert out["rel_err"] <= 1e-12, out
    assert out.get("rtile_pass_mfma", 0) >= 1, out


def test_register_resident_auto_choice_on_cpu(drive, orc, tmp_path):
    # by default a dense pass that needs no remap runs register-resident and a
    # pass that would need remaps runs on the LDS tile kernel
    rng = np.random.default_rng(66)
    n = 14
    dense = [("manyTargGate", ([4, 9, 12], _unitary(rng, 8))), ("oneTargGate", (10, _unitary(rng, 2)))]
    path = tmp_path / "auto_dense.bin"
    _circuit_file(path, orc, n, dense, rng)
    out = _run(drive, ["circuit", path])
    assert out["rel_err"] <= 1e-12 and out.get("rtile_pass_mfma", 0) == 1, out
    wide = [("oneTargGate", (q, _unitary(rng, 2))) for q in (3, 4, 5, 6, 7, 8, 9, 10, 13)]
    path = tmp_path / "auto_wide.bin"
    _circuit_file(path, orc, n, wide, rng)
    out = _run(drive, ["circuit", path])
    assert out["rel_err"] <= 1e-12 and out.get("tile_pass_plain", 0) == 1, out
