import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.load()


@pytest.fixture(scope="session")
def dev():
    """The device library.  A GPU test fails (not skips) when it is missing."""
    import sdr_amd

    lib = sdr_amd.load_library()
    n, arch = sdr_amd.device_info()
    assert n >= 1, "no HIP device visible"
    assert arch.startswith("gfx950"), arch
    return lib
