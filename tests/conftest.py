import os
import sys

import pytest

# torch (used by the device-pointer tests) ships its own HIP runtime; load it
# before libhora_mi355x.so so that one runtime serves the whole process.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build()
    return oracle
