import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libkspn.so")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (checker only)."""
    from oracle import oracle

    if not os.path.exists(oracle.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return oracle


@pytest.fixture(scope="session")
def ka():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import kaspan_amd

    return kaspan_amd
