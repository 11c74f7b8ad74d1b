import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def cuda():
    """torch with a visible GPU; a gpu-marked test without one is an error, not a skip."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("gpu test run without a visible GPU")
    torch.cuda.init()
    return torch
