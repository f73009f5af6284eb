import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def ctx():
    """One evaluator context for the GPU tests (fails loudly without the HIP lib)."""
    from symbolic_regression_gpu_amd import Context
    c = Context(0)
    yield c
    c.close()
