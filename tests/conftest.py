import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "tensor-net_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a GPU test on a machine without a device fails loudly."""
    from tnmps import _lib
    lib = _lib.load()
    n = lib.tn_device_count()
    assert n > 0, "no HIP device visible: the gpu tests need an MI355X"
    return lib
