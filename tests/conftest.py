import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("lasp-d3_amd/src", "lasp-d3_amd/tables", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device query only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def table():
    import d3_lib
    return d3_lib.table()
