import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libamgs.so")
    config.addinivalue_line("markers", "slow: longer CPU-only checks")


def pytest_collection_modifyitems(config, items):
    # GPU tests are skipped, not failed, where no device exists; on a GPU box
    # they must run and the library must load (see tests/gpu_util.py).
    try:
        import torch  # noqa: F401
        have_gpu = torch.cuda.is_available()
    except Exception:
        have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
