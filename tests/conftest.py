"""Test configuration: import paths for the product package (igl), the
workload generators and the oracle, plus the `gpu` marker."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for path in (os.path.join(ROOT, "mojo-libigl_amd", "python"), ROOT):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(autouse=True)
def _gpu_tests_never_use_the_cpu_fallback(request, monkeypatch):
    """A GPU test that asks igl for device="gpu" must run the HIP path: with
    no usable device the package's CPU fallback raises instead (the
    WN_GPU_STRICT switch), so such a test fails rather than passing on the
    CPU."""
    if request.node.get_closest_marker("gpu") is not None:
        monkeypatch.setenv("WN_GPU_STRICT", "1")


def gpu_available() -> bool:
    try:
        import igl

        return igl.lib().wn_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible")
