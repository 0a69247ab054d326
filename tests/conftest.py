import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs on the MI355X box)")


@pytest.fixture(scope="session")
def ctx():
    from tetwild_amd.ops import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda", 0)
