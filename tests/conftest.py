import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests" / "golden"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not pyoracle.LIB_PATH.exists():
        pyoracle.build()
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def progress(msg: str) -> None:
    """Flushed one-line progress note (names the case before GPU work)."""
    print(msg, flush=True)
    sys.stdout.flush()
    try:
        os.fsync(sys.stdout.fileno())
    except OSError:
        pass
