import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("omm-bundle_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _have_gpu():
    try:
        import ommhip

        return ommhip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    import ommhip

    if not _have_gpu():
        pytest.fail("GPU test selected but no HIP device is visible (or libommhip.so is not built)")
    c = ommhip.Ctx(0)
    yield c
    c.close()
