import os
import sys

import pytest

# HIP places a process's streams round-robin on GPU_MAX_HW_QUEUES hardware
# queues (4 by default); the serving tests keep four slot streams beside the
# default stream, and slots sharing a queue run their batches one after the
# other.  Raised (never lowered) before anything initialises HIP.
try:
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
except ValueError:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a gfx950 GPU and the built libd3d_hip.so")


@pytest.fixture(scope="session")
def gpu():
    """Skips (never silently falls back) when no GPU is visible; fails when
    a GPU is visible but the HIP library is not built."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from distance3d_amd import _native
    _native.lib()
    return torch.device("cuda")
