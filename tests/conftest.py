import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the GPU)")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def gpu():
    # The device is queried through libturbobfs itself: no second HIP runtime
    # (such as the one bundled with torch) is brought into the test process.
    import turbobfs_amd

    turbobfs_amd.load()  # raises when the library is missing: no fallback
    if turbobfs_amd.device_count() == 0:
        pytest.skip("no GPU")
    return turbobfs_amd.device_info(0)["name"]
