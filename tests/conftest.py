import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests (the HIP library must be built)."""
    from gazelle_plugin_amd.context import Context

    c = Context(0)
    yield c
    c.close()
