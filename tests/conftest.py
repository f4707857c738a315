import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def gpu():
    """Initialise HIP device 0 through the C ABI; fails loudly without it."""
    from kuafu_amd import lib
    from kuafu_amd.abi import check
    check(lib().kf_init(0))
    return True
