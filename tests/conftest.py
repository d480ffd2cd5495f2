import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: runs HIP kernels through libworld_amd.so on an MI355X")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import cpu

    cpu.lib()
    return cpu
