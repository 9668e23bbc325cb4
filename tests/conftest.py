import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, on a machine with a GPU."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test on a machine without a HIP device")
    import gasoline_amd as gg
    gg.lib()
    return gg
