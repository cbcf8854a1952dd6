import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) device; runs the HIP path")


@pytest.fixture(scope="session")
def gpu():
    from graphics_amd import sphgpu
    if not sphgpu.device_available():
        pytest.fail("no gfx9 HIP device visible: GPU tests must run on the MI355X")
    return True
