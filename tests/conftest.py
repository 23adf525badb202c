import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def ctx():
    """One library context for the GPU tests (fails loudly without a GPU)."""
    import cpu_gpu_coprocess_groupby_agg_amd as tga
    c = tga.Context()
    yield c
    c.close()
