import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # Build the HIP library once if it is missing and a compiler is present
    # (a GPU box runs the prebuilt library and has nothing to build).
    from wavedg_amd import capi

    if not os.path.exists(capi.LIB_PATH) and os.path.exists("/opt/rocm/bin/hipcc"):
        from wavedg_amd import _build

        _build.build()


@pytest.fixture(scope="session")
def cuda():
    """torch with a visible GPU; a gpu-marked test without one is an error, not a skip."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("gpu test run without a visible GPU")
    torch.cuda.init()
    return torch
