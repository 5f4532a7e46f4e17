import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhydrant_hip on cuda:0)")
    config.addinivalue_line("markers", "slow: full-size parity case (seconds to a minute)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from hydrant_amd.hip import Context

    ctx = Context(0)
    yield ctx
    ctx.close()
