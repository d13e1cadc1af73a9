import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size parity cases (minutes of CPU oracle time)")


@pytest.fixture(scope="session")
def orc():
    """The oracle package, with liborc.so built (restatement only)."""
    import oracle
    from oracle import oracle as mod
    if not os.path.exists(mod._ORC_PATH):
        oracle.build_oracle(ref=False)
    return oracle
