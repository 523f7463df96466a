import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def _gpu_present():
    try:
        import stardust_amd as sd
        return sd.device_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a real device; GPU tests fail loudly without it."""
    import stardust_amd as sd
    lib = sd.lib()  # raises if the library is missing
    if not sd.device_available():
        pytest.fail("no HIP device visible for a test marked gpu")
    return lib


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
