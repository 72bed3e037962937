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


def _gpu_present():
    try:
        from fftcorr_amd import _lib
        return _lib.device_count() > 0
    except RuntimeError as e:
        if "not built" in str(e):
            raise  # the library must exist wherever GPU tests run
        return False


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items or _gpu_present():
        return
    skip = pytest.mark.skip(reason="no GPU visible to HIP")
    for it in gpu_items:
        it.add_marker(skip)


@pytest.fixture(scope="session")
def fclib():
    from fftcorr_amd import _lib
    return _lib
