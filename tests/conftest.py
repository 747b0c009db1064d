import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available() -> bool:
    # through the engine's own runtime: loading a second HIP runtime (a framework's) in the same
    # process would take the device away from the engine
    import gsofatpds_amd
    return gsofatpds_amd.gpu_present()


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
