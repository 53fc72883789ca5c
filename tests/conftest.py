import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")


@pytest.fixture(scope="session")
def dev():
    """The one device context of the session (core/gpu shared_context())."""
    import torch
    from revenant_amd.ops import Device

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    d = Device(int(os.environ.get("REVENANT_GPU_INDEX", "0")))
    yield d
    d.close()
