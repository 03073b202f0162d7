"""The reference-style C++ tests (tests/cpp/test_dropin.cpp) over the drop-in
headers: every public function of the reference's API, 200 random trials each,
executed by libfsa on the GPU and compared with the oracle."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "test_dropin"


def test_dropin_headers_match_oracle():
    assert BIN.exists(), "tests/cpp/test_dropin is not built (__graft_entry__.build())"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
