import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libnmr_hip.so")
    # The oracle is plain C; build it when it is missing (cheap, CPU only).
    lib = os.path.join(ROOT, "oracle", "libnmr_oracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly when it is missing."""
    from initializator_amd import _lib

    return _lib.load()
