"""CPU-side checks of the C ABI (no GPU needed).

libfsa.so must load and export every symbol include/fsa.h declares, and
argument validation must reject bad calls before any device work.
"""
import ctypes as C
import re
from pathlib import Path

import pytest

import __graft_entry__ as ge

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(fsa_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def fsa():
    mod = ge.load_package()
    if not mod.LIB_PATH.exists():
        pytest.skip("libfsa.so not built")
    return mod


def test_header_declares_the_api():
    syms = declared_symbols()
    assert len(syms) > 40
    for must in ("fsa_statevector_oneTargGate", "fsa_densitymatrix_partialTrace", "fsa_comm_getRank"):
        assert must in syms


def test_library_exports_every_declared_symbol(fsa):
    lib = fsa.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_single_rank_communication(fsa):
    assert fsa.comm_getRank() == 0
    assert fsa.comm_getNumNodes() == 1


def test_null_state_is_rejected(fsa):
    lib = fsa.lib()
    g = (C.c_double * 8)()
    assert lib.fsa_statevector_oneTargGate(None, 0, g) == 1
    assert b"null" in lib.fsa_last_error()


def test_bad_precision_rejected_before_allocation(fsa):
    h = C.c_void_p()
    assert fsa.lib().fsa_state_create(3, 0, 7, C.byref(h)) == 1
    assert h.value is None
