import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_addoption(parser):
    parser.addoption("--dry-gpu", action="store_true",
                     help="run the gpu-marked tests on the CPU with the oracle standing in for the device "
                          "(checks the tests themselves, not the kernels)")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    if config.getoption("--dry-gpu"):
        os.environ["ENL_TEST_DRY_RUN"] = "1"


@pytest.fixture(scope="session")
def gas():
    from enlightning_amd.gas import Gas
    return Gas()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if config.getoption("--dry-gpu") or _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
