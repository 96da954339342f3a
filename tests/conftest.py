import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
# Choices the tests measure (a quick autotune, batched plans on small batches) stay in
# memory: writing them to the persistent tuning cache would let later runs of the same
# tree use schedules measured on test-sized images.
os.environ.setdefault("DIPAMD_TUNE_WRITE", "0")
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test started without a HIP device")
    return torch.device("cuda:0")
