import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "an5d-artifact_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def torch_cuda():
    """torch with its GPU.  Only tests marked `gpu` use it, so a torch that
    sees no device is a failure, not a skip (it would hide the device-tensor
    entry points)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("torch sees no GPU in this process")
    return torch
