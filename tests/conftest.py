import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fast-er_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libfaster_hip.so")
    config.addinivalue_line("markers", "slow: longer-running check")


@pytest.fixture(scope="session")
def oracle():
    import faster_oracle

    faster_oracle.lib()
    return faster_oracle
