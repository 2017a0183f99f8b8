import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qaintensor.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs through libqtn on cuda:0)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """The process-wide libqtn context.  GPU tests fail loudly (no skip, no
    CPU fallback) if the library or the device is missing."""
    from qaintensor import default_context
    return default_context()


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    # on a machine without a GPU, tests marked gpu are deselected only when
    # they were not explicitly requested with -m gpu
    markexpr = config.getoption("-m") or ""
    if "gpu" in markexpr and "not gpu" not in markexpr:
        return
    skip = pytest.mark.skip(reason="no GPU on this machine")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
