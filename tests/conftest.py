import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP stepper)")


@pytest.fixture(scope="session")
def gpu_device():
    from fdtd2d_amd._lib import device_count

    if device_count() < 1:
        pytest.skip("no GPU")
    return 0
