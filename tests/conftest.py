import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """Device description from the HIP library; fails (never skips) without a GPU."""
    from otfwi_amd import _lib

    try:
        info = _lib.device_info()
    except Exception as e:  # pragma: no cover - only without a GPU
        pytest.fail(f"GPU test run without a usable GPU: {e}")
    print("device:", info, flush=True)
    return info
