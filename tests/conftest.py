import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "funicular.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def backend():
    """The HIP backend on device 0. GPU tests use it; with no device it fails
    loudly rather than falling back to anything."""
    import funicular_hip
    return funicular_hip.hip_backend()
