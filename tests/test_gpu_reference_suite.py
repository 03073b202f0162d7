"""The reference's own Catch2 suite (tests/tests_statevector.hpp,
tests/tests_densitymatrix.hpp) and its main.cpp, unchanged, compiled against
the drop-in headers and linked to libfsa (oracle/Makefile, target ``suite``).
Every test case runs its 5000 random trials on the GPU and checks them with
the reference harness's dense matrices.  The binaries are built where the
reference sources are present and travel with the tree; without them the
tests skip."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref"
SUITE = REF / "ref_tests"
MAIN = REF / "ref_main"

# TEST_CASE names of tests/tests_statevector.hpp and tests/tests_densitymatrix.hpp
CASES = [
    "statevector_oneTargGate", "statevector_manyCtrlOneTargGate", "statevector_swapGate",
    "statevector_manyTargGate", "statevector_pauliTensor", "statevector_pauliGadget",
    "statevector_phaseGadget",
    "densitymatrix_manyTargGate", "densitymatrix_swapGate", "densitymatrix_pauliTensor",
    "densitymatrix_pauliGadget", "densitymatrix_phaseGadget", "densitymatrix_krausMap",
    "densitymatrix_oneQubitDephasing", "densitymatrix_twoQubitDephasing",
    "densitymatrix_oneQubitDepolarising", "densitymatrix_twoQubitDepolarising",
    "densitymatrix_damping", "densitymatrix_expecPauliString", "densitymatrix_partialTrace",
]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", CASES)
def test_reference_catch2_case(case):
    if not SUITE.exists():
        pytest.skip("oracle/_ref/ref_tests not built (needs the reference sources at build time)")
    print(f"reference suite case {case}", flush=True)
    r = subprocess.run([str(SUITE), case, "--rng-seed", "20261203"], capture_output=True, text=True, timeout=540)
    tail = r.stdout[-3000:] + r.stderr[-2000:]
    print(tail)
    assert r.returncode == 0, tail
    assert "All tests passed" in r.stdout, tail


@pytest.mark.timeout(300)
def test_reference_main():
    if not MAIN.exists():
        pytest.skip("oracle/_ref/ref_main not built (needs the reference sources at build time)")
    r = subprocess.run([str(MAIN)], capture_output=True, text=True, timeout=240)
    print(r.stdout[-1000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "done in" in r.stdout
