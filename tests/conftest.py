import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context on cuda:0 for the whole GPU session."""
    import torch
    import magik_amd as M
    assert torch.cuda.is_available(), "GPU tests need a GPU; run them with -m gpu on an MI355X"
    ctx = M.Context(0)
    yield ctx
    ctx.close()
