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


@pytest.fixture(scope="session")
def host_bin():
    """Directory of the host programs (nlm_fd, nlm_mesh, nlm_dc); builds the
    host layer when it is missing (on a machine with the compilers)."""
    from nlmech_amd import build as nb
    tools = ("nlm_fd", "nlm_mesh", "nlm_dc")
    if not all((nb.BINDIR / t).exists() for t in tools):
        if not nb.LIB.exists():
            nb.build()
        else:
            nb.build_host()
    return nb.BINDIR


def run_tool(path, *args, cwd=None, check=True):
    """Runs a host program as a child process; returns (rc, stdout+stderr)."""
    import subprocess
    r = subprocess.run([str(path), *map(str, args)], cwd=cwd,
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    if check and r.returncode != 0:
        raise AssertionError(f"{path} {' '.join(map(str, args))} -> "
                             f"{r.returncode}\n{out}")
    return r.returncode, out
