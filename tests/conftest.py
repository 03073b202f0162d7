import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "oracle", ROOT / "tests", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle

    r = oracle.reference()
    if r is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    return r
