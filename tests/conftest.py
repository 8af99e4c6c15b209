import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """A device context; GPU tests fail (not skip) when no GPU or no library."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    from mcvslam_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def orb_default(gpu_ctx):
    from mcvslam_amd import ORB
    return ORB(2000, 1.2, 8, 20, 7, ctx=gpu_ctx)
