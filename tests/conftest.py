import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def ref():
    """The reference's own FFTW-free code (oracle/_ref), or skip."""
    from oracle import ref as _ref
    if _ref.load() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources at build time)")
    return _ref
