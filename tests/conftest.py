import importlib
import os
import sys

import pytest

# PyTorch-ROCm carries its own HIP runtime (libamdhip64); in one process the
# first one loaded serves both it and libp2d.so, so torch goes first whenever a
# test may use it next to the library.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library's C ABI)")


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def W():
    return importlib.import_module("2d_engine_amd.workloads")


@pytest.fixture(scope="session")
def P2D():
    """The product package; loading it fails loudly if libp2d.so is not built."""
    return importlib.import_module("2d_engine_amd")
