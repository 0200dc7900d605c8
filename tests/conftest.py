import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpichip.so")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu_lib(pkg):
    """One device context for the whole GPU session (the library state is per process)."""
    n = __import__("ctypes").c_int()
    rc = pkg.lib().pic_device_count(__import__("ctypes").byref(n))
    if rc != 0 or n.value == 0:
        pytest.fail("no HIP device visible to libpichip.so")
    return pkg.PicLibrary(0)
